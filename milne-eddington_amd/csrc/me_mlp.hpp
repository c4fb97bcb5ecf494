// me_mlp.hpp — learned initialiser (MLP 4*nl -> 256 -> 256 -> 256 -> 10).
#pragma once

#include "../../include/meinv.h"
#include "me_device.hpp"

namespace me {

struct MlpState {
    int nin = 0;
    int trained = 0;
};

inline void mlp_init(MlpState &m, int nin) { m.nin = nin; m.trained = 0; }
inline void mlp_free(MlpState &m) { m.trained = 0; }
inline int64_t mlp_num_params(const MlpState &) { return 0; }
inline me_status mlp_get_params(MlpState &, hipStream_t, float *) { return ME_ENOTREADY; }
inline me_status mlp_set_params(MlpState &, hipStream_t, const float *) { return ME_ENOTREADY; }
inline me_status mlp_grad_host(MlpState &, hipStream_t, const float *, const float *, int, float *, double *)
{
    return ME_ENOTREADY;
}
inline me_status mlp_train(MlpState &, const DevCfg &, hipStream_t, uint64_t, const double *, double, double, double,
                           int64_t, int, int, double, double *)
{
    return ME_ENOTREADY;
}
inline me_status mlp_predict(MlpState &, const DevCfg &, hipStream_t, const float *, int, int64_t, double *)
{
    return ME_ENOTREADY;
}

}  // namespace me
