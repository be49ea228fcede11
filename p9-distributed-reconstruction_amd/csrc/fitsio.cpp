// FITS readers and writers for measurement sets and images (host code).
//
// Only the primary HDU is read or written, as the reference's FitsIO does
// (Core/FitsIO.cs: every reader takes f.ReadHDU() as an ImageHDU).  A file is
// memory-mapped; the header is parsed from its 80-character cards and the
// big-endian primary array is decoded straight into the caller's arrays on a
// few host threads, without the reference's nested jagged arrays.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "p9fits.h"
#include "p9idg_internal.h"

using p9idg::fail;

namespace {

constexpr int64_t kRecord = 2880;  // FITS logical record
constexpr int kCard = 80;          // header card

std::atomic<int> g_threads{0};

int reader_threads() {
    int t = g_threads.load();
    if (t > 0) return t;
    unsigned hw = std::thread::hardware_concurrency();
    return (int)std::min<unsigned>(std::max(hw, 1u), 64u);
}

// Runs body(begin, end) over [0, n) split into contiguous chunks, one per
// thread; small ranges stay on the calling thread.
template <typename F>
void parallel_for(int64_t n, int64_t min_per_thread, F body) {
    int nt = (int)std::min<int64_t>(reader_threads(), std::max<int64_t>(n / std::max<int64_t>(min_per_thread, 1), 1));
    if (nt <= 1) {
        body(int64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    int64_t chunk = (n + nt - 1) / nt;
    for (int i = 1; i < nt; ++i) {
        int64_t b = i * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([=] { body(b, e); });
    }
    body(0, std::min(n, chunk));
    for (auto& th : pool) th.join();
}

inline double be_f64(const uint8_t* p) {
    uint64_t u;
    std::memcpy(&u, p, 8);
    u = __builtin_bswap64(u);
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}

inline float be_f32(const uint8_t* p) {
    uint32_t u;
    std::memcpy(&u, p, 4);
    u = __builtin_bswap32(u);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

inline void put_be64(uint8_t* p, double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    u = __builtin_bswap64(u);
    std::memcpy(p, &u, 8);
}

inline void put_be32(uint8_t* p, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u = __builtin_bswap32(u);
    std::memcpy(p, &u, 4);
}

// A read-only mapping of a whole file.
struct Mapping {
    int fd = -1;
    const uint8_t* data = nullptr;
    size_t size = 0;
    ~Mapping() {
        if (data) munmap(const_cast<uint8_t*>(data), size);
        if (fd >= 0) close(fd);
    }
};

int open_mapped(const char* path, Mapping& m) {
    if (!path) return fail(P9IDG_ERR_ARGUMENT, "path is NULL");
    m.fd = open(path, O_RDONLY | O_CLOEXEC);
    if (m.fd < 0) return fail(P9IDG_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    struct stat st;
    if (fstat(m.fd, &st) != 0) return fail(P9IDG_ERR_IO, std::string("cannot stat ") + path + ": " + std::strerror(errno));
    m.size = (size_t)st.st_size;
    if (m.size < (size_t)kRecord) return fail(P9IDG_ERR_IO, std::string(path) + " is not a FITS file (shorter than one record)");
    void* p = mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (p == MAP_FAILED) return fail(P9IDG_ERR_IO, std::string("cannot map ") + path + ": " + std::strerror(errno));
    m.data = static_cast<const uint8_t*>(p);
    madvise(p, m.size, MADV_SEQUENTIAL);
    return P9IDG_OK;
}

std::string keyword(const char* card) {
    std::string k(card, 8);
    k.erase(k.find_last_not_of(' ') + 1);
    return k;
}

// Value field of a "KEYWORD = value / comment" card (columns 11-80), without
// the comment and surrounding blanks; false when the card has no value
// indicator in columns 9-10.
bool card_value(const char* card, std::string& v) {
    if (card[8] != '=' || card[9] != ' ') return false;
    std::string s(card + 10, kCard - 10);
    size_t slash = s.find('/');
    if (slash != std::string::npos) s.resize(slash);
    size_t b = s.find_first_not_of(' ');
    if (b == std::string::npos) {
        v.clear();
        return true;
    }
    size_t e = s.find_last_not_of(' ');
    v = s.substr(b, e - b + 1);
    return true;
}

bool parse_int(const std::string& v, int64_t& out) {
    if (v.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (v[0] == '+' || v[0] == '-') {
        neg = v[0] == '-';
        i = 1;
    }
    if (i >= v.size()) return false;
    int64_t x = 0;
    for (; i < v.size(); ++i) {
        if (v[i] < '0' || v[i] > '9') return false;
        if (x > (INT64_MAX - 9) / 10) return false;
        x = x * 10 + (v[i] - '0');
    }
    out = neg ? -x : x;
    return true;
}

// Parses the primary header of a mapped file (FITS Standard 4.0, section 4):
// SIMPLE = T on the first card, then BITPIX, NAXIS and NAXIS1..NAXISn, up to
// the END card; the data start at the next record boundary.
int parse_header(const char* path, const Mapping& m, p9fits_info& info) {
    std::memset(&info, 0, sizeof(info));
    const std::string where = std::string(path) + ": ";
    const char* first = reinterpret_cast<const char*>(m.data);
    std::string v;
    if (keyword(first) != "SIMPLE" || !card_value(first, v) || v != "T")
        return fail(P9IDG_ERR_IO, where + "not a FITS file (first card is not SIMPLE = T)");
    bool have_bitpix = false, have_naxis = false;
    int64_t naxes_seen = 0;
    int64_t end_card = -1;
    const int64_t ncards = (int64_t)(m.size / kCard);
    for (int64_t i = 1; i < ncards; ++i) {
        const char* card = first + i * kCard;
        std::string k = keyword(card);
        if (k == "END") {
            end_card = i;
            break;
        }
        if (!card_value(card, v)) continue;
        int64_t x;
        if (k == "BITPIX") {
            if (!parse_int(v, x)) return fail(P9IDG_ERR_IO, where + "BITPIX is not an integer");
            if (x != 8 && x != 16 && x != 32 && x != 64 && x != -32 && x != -64)
                return fail(P9IDG_ERR_IO, where + "BITPIX " + std::to_string(x) + " is not a FITS data type");
            info.bitpix = (int32_t)x;
            have_bitpix = true;
        } else if (k == "NAXIS") {
            if (!parse_int(v, x) || x < 0 || x > 999) return fail(P9IDG_ERR_IO, where + "bad NAXIS");
            if (x > P9FITS_MAX_AXES)
                return fail(P9IDG_ERR_IO, where + "NAXIS " + std::to_string(x) + " exceeds " +
                                              std::to_string(P9FITS_MAX_AXES) + " axes");
            info.naxis = (int32_t)x;
            have_naxis = true;
        } else if (k.size() > 5 && k.compare(0, 5, "NAXIS") == 0) {
            int64_t n;
            if (!parse_int(k.substr(5), n) || n < 1) continue;
            if (!have_naxis || n > info.naxis) continue;  // beyond NAXIS: not a data axis
            if (!parse_int(v, x) || x < 0) return fail(P9IDG_ERR_IO, where + k + " is not a non-negative integer");
            info.naxes[n - 1] = x;
            naxes_seen |= int64_t(1) << (n - 1);
        }
    }
    if (end_card < 0) return fail(P9IDG_ERR_IO, where + "no END card in the header");
    if (!have_bitpix || !have_naxis) return fail(P9IDG_ERR_IO, where + "BITPIX or NAXIS missing");
    if (naxes_seen != (int64_t(1) << info.naxis) - 1)
        return fail(P9IDG_ERR_IO, where + "NAXISn keywords missing for some of the " + std::to_string(info.naxis) + " axes");
    info.data_offset = ((end_card + 1) * kCard + kRecord - 1) / kRecord * kRecord;
    __int128 bytes = info.naxis ? std::abs(info.bitpix) / 8 : 0;
    for (int a = 0; a < info.naxis; ++a) {
        bytes *= info.naxes[a];
        if (bytes > ((__int128)1 << 62)) return fail(P9IDG_ERR_IO, where + "primary array too large");
    }
    info.data_bytes = (int64_t)bytes;
    if ((__int128)info.data_offset + bytes > (__int128)m.size)
        return fail(P9IDG_ERR_IO, where + "file is truncated: the header announces " + std::to_string(info.data_bytes) +
                                      " data bytes, " + std::to_string((int64_t)m.size - info.data_offset) +
                                      " are present");
    return P9IDG_OK;
}

struct Source {
    Mapping map;
    p9fits_info info;
    const uint8_t* data() const { return map.data + info.data_offset; }
};

std::string shape_text(const p9fits_info& f) {
    std::string s = "BITPIX " + std::to_string(f.bitpix) + ", NAXIS " + std::to_string(f.naxis) + " (";
    for (int a = 0; a < f.naxis; ++a) s += (a ? " x " : "") + std::to_string(f.naxes[a]);
    return s + ")";
}

// Opens `path` and checks that its primary array has the element type and
// number of axes the reader's casts require.
int open_array(const char* path, int bitpix, int naxis, const char* what, Source& src) {
    int st = open_mapped(path, src.map);
    if (st) return st;
    st = parse_header(path, src.map, src.info);
    if (st) return st;
    if (src.info.bitpix != bitpix || src.info.naxis != naxis)
        return fail(P9IDG_ERR_IO, std::string(path) + ": " + what + " needs BITPIX " + std::to_string(bitpix) +
                                      " with NAXIS " + std::to_string(naxis) + ", the file has " + shape_text(src.info));
    return P9IDG_OK;
}

// Resolves [from_bl, to_bl) against the file's baselines (to_bl = -1: all).
int baseline_range(const char* path, int64_t nbl, int32_t from_bl, int32_t& to_bl) {
    if (to_bl == -1) to_bl = (int32_t)std::min<int64_t>(nbl, INT32_MAX);
    if (from_bl < 0 || to_bl < from_bl || to_bl > nbl)
        return fail(P9IDG_ERR_ARGUMENT, std::string(path) + ": baseline range [" + std::to_string(from_bl) + ", " +
                                            std::to_string(to_bl) + ") is outside the file's " + std::to_string(nbl) +
                                            " baselines");
    return P9IDG_OK;
}

int check_leading(const char* path, const char* axis, int64_t want, int64_t have) {
    if (want < 0 || want > have)
        return fail(P9IDG_ERR_ARGUMENT, std::string(path) + ": " + std::to_string(want) + " " + axis +
                                            " requested, the file holds " + std::to_string(have));
    return P9IDG_OK;
}

}  // namespace

extern "C" {

int p9fits_set_threads(int32_t threads) {
    g_threads.store(threads > 0 ? std::min(threads, 1024) : 0);
    return P9IDG_OK;
}

int p9fits_read_info(const char* path, p9fits_info* out) {
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "out is NULL");
    Mapping m;
    int st = open_mapped(path, m);
    if (st) return st;
    return parse_header(path, m, *out);
}

int p9fits_read_frequencies(const char* path, double* out, int64_t capacity, int64_t* count) {
    Source src;
    int st = open_array(path, -64, 1, "ReadFrequencies", src);
    if (st) return st;
    const int64_t n = src.info.naxes[0];
    if (count) *count = n;
    if (!out) return P9IDG_OK;
    if (capacity < n)
        return fail(P9IDG_ERR_CAPACITY, std::string(path) + ": " + std::to_string(n) + " frequencies, capacity " +
                                            std::to_string(capacity));
    const uint8_t* d = src.data();
    for (int64_t i = 0; i < n; ++i) out[i] = be_f64(d + 8 * i);
    return P9IDG_OK;
}

int p9fits_count_baselines(const char* path, int32_t* baselines, int32_t* timesteps) {
    Source src;
    int st = open_array(path, -64, 3, "ReadUVW", src);
    if (st) return st;
    if (src.info.naxes[2] > INT32_MAX || src.info.naxes[1] > INT32_MAX)
        return fail(P9IDG_ERR_IO, std::string(path) + ": more than 2^31 baselines or timesteps");
    if (baselines) *baselines = (int32_t)src.info.naxes[2];
    if (timesteps) *timesteps = (int32_t)src.info.naxes[1];
    return P9IDG_OK;
}

int p9fits_read_uvw(const char* path, int32_t from_bl, int32_t to_bl, double* uvw) {
    Source src;
    int st = open_array(path, -64, 3, "ReadUVW", src);
    if (st) return st;
    const p9fits_info& f = src.info;
    if (f.naxes[0] < 3) return fail(P9IDG_ERR_IO, std::string(path) + ": uvw axis (NAXIS1) shorter than 3");
    if ((st = baseline_range(path, f.naxes[2], from_bl, to_bl))) return st;
    if (!uvw && to_bl > from_bl) return fail(P9IDG_ERR_ARGUMENT, "uvw is NULL");
    const int64_t T = f.naxes[1], N1 = f.naxes[0];
    const uint8_t* d = src.data();
    const int64_t rows = (int64_t)(to_bl - from_bl) * T;  // (baseline, timestep) pairs
    parallel_for(rows, 4096, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            const uint8_t* p = d + 8 * (((int64_t)from_bl * T + r) * N1);
            uvw[3 * r + 0] = be_f64(p);
            uvw[3 * r + 1] = -be_f64(p + 8);  // FitsIO.cs:214: v is flipped
            uvw[3 * r + 2] = be_f64(p + 16);
        }
    });
    return P9IDG_OK;
}

int p9fits_read_visibilities(const char* path, int32_t from_bl, int32_t to_bl, int32_t timesteps,
                             int32_t channels, double norm, void* vis, int32_t vis_type) {
    if (vis_type != P9IDG_C128 && vis_type != P9IDG_C64)
        return fail(P9IDG_ERR_ARGUMENT, "vis_type must be P9IDG_C128 or P9IDG_C64");
    Source src;
    int st = open_array(path, -64, 5, "ReadVisibilities", src);
    if (st) return st;
    const p9fits_info& f = src.info;
    // axes: NAXIS1 complex component, NAXIS2 polarization, NAXIS3 channel,
    // NAXIS4 timestep, NAXIS5 baseline
    if (f.naxes[0] < 2 || f.naxes[1] < 4)
        return fail(P9IDG_ERR_IO, std::string(path) + ": needs 2 complex components and 4 polarizations per channel, has " +
                                      shape_text(f));
    if ((st = baseline_range(path, f.naxes[4], from_bl, to_bl))) return st;
    if ((st = check_leading(path, "timesteps", timesteps, f.naxes[3]))) return st;
    if ((st = check_leading(path, "channels", channels, f.naxes[2]))) return st;
    const int64_t nbl = to_bl - from_bl;
    if (!vis && nbl * timesteps * channels > 0) return fail(P9IDG_ERR_ARGUMENT, "vis is NULL");
    const int64_t N1 = f.naxes[0], N2 = f.naxes[1], N3 = f.naxes[2], N4 = f.naxes[3];
    const int64_t pol = N1 * N2;  // doubles per channel
    const uint8_t* d = src.data();
    double* o128 = static_cast<double*>(vis);
    float* o64 = static_cast<float*>(vis);
    const bool c64 = vis_type == P9IDG_C64;
    parallel_for(nbl * timesteps, 256, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            const int64_t bl = from_bl + r / timesteps, t = r % timesteps;
            const uint8_t* row = d + 8 * ((bl * N4 + t) * N3 * pol);
            for (int64_t c = 0; c < channels; ++c) {
                const uint8_t* xx = row + 8 * (c * pol);         // polarization 0
                const uint8_t* yy = row + 8 * (c * pol + 3 * N1);  // polarization 3
                const double re = (be_f64(xx) + be_f64(yy)) / norm;
                const double im = (be_f64(xx + 8) + be_f64(yy + 8)) / norm;
                const int64_t o = 2 * (r * channels + c);
                if (c64) {
                    o64[o] = (float)re;
                    o64[o + 1] = (float)im;
                } else {
                    o128[o] = re;
                    o128[o + 1] = im;
                }
            }
        }
    });
    return P9IDG_OK;
}

int p9fits_read_flags(const char* path, int32_t from_bl, int32_t to_bl, int32_t timesteps, int32_t channels,
                      uint8_t* flags) {
    Source src;
    int st = open_array(path, -64, 4, "ReadFlags", src);
    if (st) return st;
    const p9fits_info& f = src.info;
    // axes: NAXIS1 polarization, NAXIS2 channel, NAXIS3 timestep, NAXIS4 baseline
    if ((st = baseline_range(path, f.naxes[3], from_bl, to_bl))) return st;
    if ((st = check_leading(path, "timesteps", timesteps, f.naxes[2]))) return st;
    if ((st = check_leading(path, "channels", channels, f.naxes[1]))) return st;
    const int64_t nbl = to_bl - from_bl;
    if (!flags && nbl * timesteps * channels > 0) return fail(P9IDG_ERR_ARGUMENT, "flags is NULL");
    const int64_t P = f.naxes[0], N2 = f.naxes[1], N3 = f.naxes[2];
    const uint8_t* d = src.data();
    parallel_for(nbl * timesteps, 256, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            const int64_t bl = from_bl + r / timesteps, t = r % timesteps;
            const uint8_t* row = d + 8 * ((bl * N3 + t) * N2 * P);
            for (int64_t c = 0; c < channels; ++c) {
                double sum = 0.0;  // summed in polarization order (FitsIO.cs:344-347)
                for (int64_t p = 0; p < P; ++p) sum += be_f64(row + 8 * (c * P + p));
                flags[r * channels + c] = sum == 0.0 ? 0 : 1;
            }
        }
    });
    return P9IDG_OK;
}

int p9fits_read_image(const char* path, int32_t type, void* out, int32_t rows, int32_t cols) {
    if (type != P9FITS_F32 && type != P9FITS_F64)
        return fail(P9IDG_ERR_ARGUMENT, "p9fits_read_image reads P9FITS_F32 or P9FITS_F64");
    Source src;
    const bool f32 = type == P9FITS_F32;
    int st = open_array(path, f32 ? -32 : -64, 2, f32 ? "ReadImage" : "ReadImageDouble", src);
    if (st) return st;
    if (rows != src.info.naxes[1] || cols != src.info.naxes[0])
        return fail(P9IDG_ERR_ARGUMENT, std::string(path) + ": image is " + std::to_string(src.info.naxes[1]) + " x " +
                                            std::to_string(src.info.naxes[0]) + ", not " + std::to_string(rows) +
                                            " x " + std::to_string(cols));
    const int64_t n = (int64_t)rows * cols;
    if (!out && n) return fail(P9IDG_ERR_ARGUMENT, "out is NULL");
    const uint8_t* d = src.data();
    parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
        if (f32)
            for (int64_t i = b; i < e; ++i) static_cast<float*>(out)[i] = be_f32(d + 4 * i);
        else
            for (int64_t i = b; i < e; ++i) static_cast<double*>(out)[i] = be_f64(d + 8 * i);
    });
    return P9IDG_OK;
}

int p9fits_read_casa(const char* path, float* out, int32_t rows, int32_t cols) {
    Source src;
    int st = open_array(path, -32, 4, "ReadCASAFits", src);
    if (st) return st;
    const p9fits_info& f = src.info;
    if (f.naxes[2] < 1 || f.naxes[3] < 1) return fail(P9IDG_ERR_IO, std::string(path) + ": no image plane");
    if (rows != f.naxes[1] || cols != f.naxes[0])
        return fail(P9IDG_ERR_ARGUMENT, std::string(path) + ": image plane is " + std::to_string(f.naxes[1]) + " x " +
                                            std::to_string(f.naxes[0]) + ", not " + std::to_string(rows) + " x " +
                                            std::to_string(cols));
    const int64_t n = (int64_t)rows * cols;
    if (!out && n) return fail(P9IDG_ERR_ARGUMENT, "out is NULL");
    const uint8_t* d = src.data();
    parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) out[i] = be_f32(d + 4 * i);
    });
    return P9IDG_OK;
}

int p9fits_write_image(const char* path, int32_t type, const void* img, int32_t rows, int32_t cols) {
    if (!path) return fail(P9IDG_ERR_ARGUMENT, "path is NULL");
    if (type < P9FITS_F32 || type > P9FITS_C128_IMAG) return fail(P9IDG_ERR_ARGUMENT, "unknown p9fits_type");
    if (rows < 1 || cols < 1) return fail(P9IDG_ERR_ARGUMENT, "image must have at least one row and column");
    if (!img) return fail(P9IDG_ERR_ARGUMENT, "img is NULL");
    const int bitpix = type == P9FITS_F32 ? -32 : -64;
    const int64_t n = (int64_t)rows * cols, esize = bitpix == -32 ? 4 : 8;

    // Mandatory keywords in fixed format: keyword in columns 1-8, "= " in
    // 9-10, value right-justified to column 30 (FITS Standard 4.0, 4.2).
    std::string header;
    auto card = [&](const char* key, const std::string& value) {
        char buf[kCard + 1];
        std::snprintf(buf, sizeof(buf), "%-8s= %20s", key, value.c_str());
        std::string c(buf);
        c.resize(kCard, ' ');
        header += c;
    };
    card("SIMPLE", "T");
    card("BITPIX", std::to_string(bitpix));
    card("NAXIS", "2");
    card("NAXIS1", std::to_string(cols));
    card("NAXIS2", std::to_string(rows));
    card("EXTEND", "T");
    {
        std::string end = "END";
        end.resize(kCard, ' ');
        header += end;
    }
    header.resize((header.size() + kRecord - 1) / kRecord * kRecord, ' ');

    const int64_t data_bytes = n * esize;
    std::vector<uint8_t> data((size_t)((data_bytes + kRecord - 1) / kRecord * kRecord), 0);
    uint8_t* o = data.data();
    parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
        switch (type) {
            case P9FITS_F32:
                for (int64_t i = b; i < e; ++i) put_be32(o + 4 * i, static_cast<const float*>(img)[i]);
                break;
            case P9FITS_F64:
                for (int64_t i = b; i < e; ++i) put_be64(o + 8 * i, static_cast<const double*>(img)[i]);
                break;
            default: {
                const double* c = static_cast<const double*>(img) + (type == P9FITS_C128_IMAG ? 1 : 0);
                for (int64_t i = b; i < e; ++i) put_be64(o + 8 * i, c[2 * i]);
            }
        }
    });

    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(P9IDG_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    bool ok = std::fwrite(header.data(), 1, header.size(), fp) == header.size() &&
              std::fwrite(data.data(), 1, data.size(), fp) == data.size();
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return fail(P9IDG_ERR_IO, std::string("write to ") + path + " failed: " + std::strerror(errno));
    return P9IDG_OK;
}

}  // extern "C"
