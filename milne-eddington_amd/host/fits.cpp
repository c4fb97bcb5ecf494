// fits.cpp — see fits.hpp.
#include "fits.hpp"

#include <sys/types.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace mefits {

namespace {

constexpr int64_t kBlock = 2880;
constexpr int kCard = 80;

struct File {
    FILE *f = nullptr;
    File(const std::string &p, const char *mode) : f(std::fopen(p.c_str(), mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// value field of a "KEYWORD = value / comment" card (columns 11-80)
std::string card_value(const char *card)
{
    std::string v(card + 10, kCard - 10);
    if (!v.empty() && v[0] != '\'') {
        size_t slash = v.find('/');
        if (slash != std::string::npos) v = v.substr(0, slash);
    }
    return trim(v);
}

bool parse_int(const std::string &s, int64_t &out)
{
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    long long v = std::strtoll(s.c_str(), &end, 10);
    if (errno || !end || trim(end).size()) return false;
    out = v;
    return true;
}

bool parse_real(std::string s, double &out)
{
    for (char &c : s)
        if (c == 'D' || c == 'd') c = 'E';  // FITS allows D exponents
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    double v = std::strtod(s.c_str(), &end);
    if (errno || !end || trim(end).size()) return false;
    out = v;
    return true;
}

int64_t elem_bytes(int bitpix) { return (bitpix < 0 ? -bitpix : bitpix) / 8; }

template <typename U> U bswap(U v);
template <> uint16_t bswap(uint16_t v) { return __builtin_bswap16(v); }
template <> uint32_t bswap(uint32_t v) { return __builtin_bswap32(v); }
template <> uint64_t bswap(uint64_t v) { return __builtin_bswap64(v); }

template <typename U, typename V> V from_be(const unsigned char *p)
{
    U u;
    std::memcpy(&u, p, sizeof(U));
    u = bswap<U>(u);
    V v;
    std::memcpy(&v, &u, sizeof(V));
    return v;
}

void put_card(std::string &hdr, const std::string &key, const std::string &value, const std::string &comment)
{
    char card[kCard + 1];
    std::snprintf(card, sizeof(card), "%-8.8s= %20s / %-47.47s", key.c_str(), value.c_str(), comment.c_str());
    hdr.append(card, kCard);
}

}  // namespace

int64_t Image::npix() const
{
    if (naxes.empty()) return 0;
    int64_t n = 1;
    for (int64_t a : naxes) n *= a;
    return n;
}

bool open_image(const std::string &path, Image &img, std::string &err)
{
    File fh(path, "rb");
    if (!fh.f) {
        err = "cannot open " + path;
        return false;
    }
    img = Image();
    img.path = path;
    int64_t naxis = -1;
    bool simple = false, end = false;
    std::vector<int64_t> ax;
    char block[kBlock];
    int64_t nblocks = 0;
    while (!end) {
        if (std::fread(block, 1, kBlock, fh.f) != (size_t)kBlock) {
            err = path + ": header ends before its END card";
            return false;
        }
        nblocks++;
        for (int c = 0; c < kBlock / kCard && !end; c++) {
            const char *card = block + c * kCard;
            std::string key = trim(std::string(card, 8));
            if (nblocks == 1 && c == 0) {
                if (key != "SIMPLE" || card_value(card) != "T") {
                    err = path + ": not a FITS file (SIMPLE = T missing)";
                    return false;
                }
                simple = true;
                continue;
            }
            if (key == "END") {
                end = true;
                break;
            }
            if (card[8] != '=') continue;
            const std::string v = card_value(card);
            int64_t iv = 0;
            double dv = 0;
            if (key == "BITPIX") {
                if (!parse_int(v, iv)) { err = path + ": bad BITPIX"; return false; }
                img.bitpix = (int)iv;
            } else if (key == "NAXIS") {
                if (!parse_int(v, iv) || iv < 0 || iv > 999) { err = path + ": bad NAXIS"; return false; }
                naxis = iv;
                ax.assign((size_t)naxis, -1);
            } else if (key.size() > 5 && key.compare(0, 5, "NAXIS") == 0) {
                int64_t k = 0;
                if (!parse_int(key.substr(5), k) || k < 1 || k > (int64_t)ax.size() || !parse_int(v, iv) || iv < 0) {
                    err = path + ": bad " + key;
                    return false;
                }
                ax[(size_t)(k - 1)] = iv;
            } else if (key == "BSCALE") {
                if (!parse_real(v, dv)) { err = path + ": bad BSCALE"; return false; }
                img.bscale = dv;
            } else if (key == "BZERO") {
                if (!parse_real(v, dv)) { err = path + ": bad BZERO"; return false; }
                img.bzero = dv;
            }
        }
    }
    if (!simple || naxis < 0) {
        err = path + ": no NAXIS in the primary header";
        return false;
    }
    for (int64_t a : ax)
        if (a < 0) {
            err = path + ": missing NAXISn";
            return false;
        }
    const int b = img.bitpix;
    if (b != 8 && b != 16 && b != 32 && b != 64 && b != -32 && b != -64) {
        err = path + ": unsupported BITPIX " + std::to_string(b);
        return false;
    }
    img.naxes = ax;
    img.data_offset = nblocks * kBlock;
    return true;
}

bool read_pixels(const Image &img, int64_t first, int64_t count, double *out, std::string &err)
{
    if (first < 0 || count < 0 || first + count > img.npix()) {
        err = img.path + ": pixel range outside the image";
        return false;
    }
    if (count == 0) return true;
    File fh(img.path, "rb");
    if (!fh.f) {
        err = "cannot open " + img.path;
        return false;
    }
    const int64_t eb = elem_bytes(img.bitpix);
    if (fseeko(fh.f, (off_t)(img.data_offset + first * eb), SEEK_SET) != 0) {
        err = img.path + ": seek failed";
        return false;
    }
    const bool scaled = img.bscale != 1.0 || img.bzero != 0.0;
    std::vector<unsigned char> buf((size_t)(eb * 65536));
    int64_t done = 0;
    while (done < count) {
        const int64_t n = count - done < 65536 ? count - done : 65536;
        if (std::fread(buf.data(), (size_t)eb, (size_t)n, fh.f) != (size_t)n) {
            err = img.path + ": data shorter than its header says";
            return false;
        }
        for (int64_t i = 0; i < n; i++) {
            const unsigned char *p = buf.data() + i * eb;
            double v;
            switch (img.bitpix) {
            case 8: v = (double)p[0]; break;
            case 16: v = (double)from_be<uint16_t, int16_t>(p); break;
            case 32: v = (double)from_be<uint32_t, int32_t>(p); break;
            case 64: v = (double)from_be<uint64_t, int64_t>(p); break;
            case -32: v = (double)from_be<uint32_t, float>(p); break;
            default: v = from_be<uint64_t, double>(p); break;
            }
            out[done + i] = scaled ? img.bzero + img.bscale * v : v;
        }
        done += n;
    }
    return true;
}

bool create_image(const std::string &path, const std::vector<int64_t> &naxes, Image &img, std::string &err)
{
    std::string hdr;
    put_card(hdr, "SIMPLE", "T", "file does conform to FITS standard");
    put_card(hdr, "BITPIX", "-64", "number of bits per data pixel");
    put_card(hdr, "NAXIS", std::to_string(naxes.size()), "number of data axes");
    for (size_t k = 0; k < naxes.size(); k++)
        put_card(hdr, "NAXIS" + std::to_string(k + 1), std::to_string(naxes[k]),
                 "length of data axis " + std::to_string(k + 1));
    put_card(hdr, "EXTEND", "T", "FITS dataset may contain extensions");
    hdr.append("END");
    hdr.append((size_t)(kBlock - hdr.size() % kBlock) % kBlock, ' ');
    int64_t npix = 1;
    for (int64_t a : naxes) npix *= a;
    int64_t data = npix * 8;
    data += (kBlock - data % kBlock) % kBlock;
    {
        File fh(path, "wb");
        if (!fh.f) {
            err = "cannot create " + path;
            return false;
        }
        if (std::fwrite(hdr.data(), 1, hdr.size(), fh.f) != hdr.size()) {
            err = path + ": write failed";
            return false;
        }
        std::fflush(fh.f);
        // the data area reads as zeros (IEEE +0.0) once the file is extended
        if (ftruncate(fileno(fh.f), (off_t)(hdr.size() + data)) != 0) {
            err = path + ": cannot size the data area";
            return false;
        }
    }
    img = Image();
    img.path = path;
    img.bitpix = -64;
    img.naxes = naxes;
    img.data_offset = (int64_t)hdr.size();
    return true;
}

bool write_pixels(const Image &img, int64_t first, int64_t count, const double *in, std::string &err)
{
    if (img.bitpix != -64) {
        err = img.path + ": results are written to BITPIX -64 images only";
        return false;
    }
    if (first < 0 || count < 0 || first + count > img.npix()) {
        err = img.path + ": pixel range outside the image";
        return false;
    }
    if (count == 0) return true;
    File fh(img.path, "r+b");
    if (!fh.f) {
        err = "cannot open " + img.path + " for writing";
        return false;
    }
    if (fseeko(fh.f, (off_t)(img.data_offset + first * 8), SEEK_SET) != 0) {
        err = img.path + ": seek failed";
        return false;
    }
    std::vector<uint64_t> buf(65536);
    int64_t done = 0;
    while (done < count) {
        const int64_t n = count - done < 65536 ? count - done : 65536;
        for (int64_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, in + done + i, 8);
            buf[(size_t)i] = __builtin_bswap64(u);
        }
        if (std::fwrite(buf.data(), 8, (size_t)n, fh.f) != (size_t)n) {
            err = img.path + ": write failed";
            return false;
        }
        done += n;
    }
    return true;
}

}  // namespace mefits
