// cb_h5.cpp -- the HDF5 snapshot file writer (Bubble::OutField, src/bubble.cpp:
// 887-954): datasets phi, pi [, Ekin, gradphi, potEnergy] as float with the
// shape of the saved points, chunked by rows, shuffle + deflate filters, and
// the double attributes D, L, nGrids, reduc, c, t, a on the root group.
//
// The HDF5 C library (1.10.3 or later, for H5Dwrite_chunk) is loaded at run
// time, so the rest of the library works where it is absent: $CB_HDF5_LIBRARY,
// then libhdf5.so on the loader's path, then the usual install locations.
// The filters run here rather than inside HDF5: the chunk images arrive
// already shuffled (on the device, or by cb_write_snapshot_h5 below), host
// threads deflate the chunks of a dataset in parallel with zlib's compress2
// (the stream HDF5's deflate filter writes), and H5Dwrite_chunk stores them.
// A reader sees the same file as one written through HDF5's own pipeline.
#include <dlfcn.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cb_sim.hpp"
#include "cb_snapshot.hpp"

using cb::set_error;

namespace {

// HDF5 1.10 C ABI: hid_t is 64-bit, herr_t / htri_t int, hsize_t 64-bit.
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr unsigned kAccTrunc = 0x0002u;   // H5F_ACC_TRUNC
constexpr hid_t kDefault = 0;             // H5P_DEFAULT
constexpr int kScalar = 0;                // H5S_SCALAR
constexpr int kFilterDeflate = 1;         // H5Z_FILTER_DEFLATE

struct H5Api {
    bool tried = false, ok = false;
    std::string why;
    void *lib = nullptr;
    herr_t (*open)();
    hid_t (*Fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Screate_simple)(int, const hsize_t *, const hsize_t *);
    hid_t (*Screate)(int);
    herr_t (*Sclose)(hid_t);
    hid_t (*Pcreate)(hid_t);
    herr_t (*Pset_chunk)(hid_t, int, const hsize_t *);
    herr_t (*Pset_shuffle)(hid_t);
    herr_t (*Pset_deflate)(hid_t, unsigned);
    herr_t (*Pclose)(hid_t);
    hid_t (*Dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dwrite_chunk)(hid_t, hid_t, uint32_t, const hsize_t *, size_t, const void *);
    herr_t (*Dclose)(hid_t);
    hid_t (*Acreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Awrite)(hid_t, hid_t, const void *);
    herr_t (*Aclose)(hid_t);
    htri_t (*Zfilter_avail)(int);
    const hid_t *native_float = nullptr, *native_double = nullptr, *dataset_create = nullptr;
};

H5Api g_h5;
std::mutex g_h5_mu;

template <typename F>
bool sym(void *lib, const char *name, F &f, std::string &missing)
{
    void *p = dlsym(lib, name);
    if (!p) {
        missing += std::string(missing.empty() ? "" : ", ") + name;
        return false;
    }
    f = reinterpret_cast<F>(p);
    return true;
}

bool load_h5(std::string &why)
{
    std::lock_guard<std::mutex> lk(g_h5_mu);
    H5Api &h = g_h5;
    if (h.tried) {
        why = h.why;
        return h.ok;
    }
    h.tried = true;
    std::vector<std::string> cands;
    if (const char *e = std::getenv("CB_HDF5_LIBRARY")) cands.push_back(e);
    for (const char *c : {"libhdf5.so", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so",
                          "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so", "/usr/local/lib/libhdf5.so"})
        cands.push_back(c);
    std::string tried;
    for (const std::string &c : cands) {
        h.lib = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (h.lib) break;
        tried += "\n  " + c + ": " + dlerror();
    }
    if (!h.lib) {
        h.why = "the HDF5 C library was not found (set CB_HDF5_LIBRARY to its path, or turn the field "
                "snapshots off); tried:" + tried;
        why = h.why;
        return false;
    }
    std::string missing;
    sym(h.lib, "H5open", h.open, missing);
    sym(h.lib, "H5Fcreate", h.Fcreate, missing);
    sym(h.lib, "H5Fclose", h.Fclose, missing);
    sym(h.lib, "H5Screate_simple", h.Screate_simple, missing);
    sym(h.lib, "H5Screate", h.Screate, missing);
    sym(h.lib, "H5Sclose", h.Sclose, missing);
    sym(h.lib, "H5Pcreate", h.Pcreate, missing);
    sym(h.lib, "H5Pset_chunk", h.Pset_chunk, missing);
    sym(h.lib, "H5Pset_shuffle", h.Pset_shuffle, missing);
    sym(h.lib, "H5Pset_deflate", h.Pset_deflate, missing);
    sym(h.lib, "H5Pclose", h.Pclose, missing);
    sym(h.lib, "H5Dcreate2", h.Dcreate2, missing);
    sym(h.lib, "H5Dwrite_chunk", h.Dwrite_chunk, missing);
    sym(h.lib, "H5Dclose", h.Dclose, missing);
    sym(h.lib, "H5Acreate2", h.Acreate2, missing);
    sym(h.lib, "H5Awrite", h.Awrite, missing);
    sym(h.lib, "H5Aclose", h.Aclose, missing);
    sym(h.lib, "H5Zfilter_avail", h.Zfilter_avail, missing);
    sym(h.lib, "H5T_NATIVE_FLOAT_g", h.native_float, missing);
    sym(h.lib, "H5T_NATIVE_DOUBLE_g", h.native_double, missing);
    sym(h.lib, "H5P_CLS_DATASET_CREATE_ID_g", h.dataset_create, missing);
    if (!missing.empty()) {
        h.why = "the HDF5 library lacks " + missing + " (version 1.10.3 or later is needed)";
    } else if (h.open() < 0) {
        h.why = "H5open failed";
    } else if (h.Zfilter_avail(kFilterDeflate) <= 0) {
        h.why = "the HDF5 library has no deflate filter";
    } else {
        h.ok = true;
    }
    why = h.why;
    return h.ok;
}

// Runs f(i) for i in [0, n) on up to `threads` host threads.
template <typename F>
void parallel_for(long long n, int threads, F f)
{
    threads = (int)std::max(1LL, std::min<long long>(threads, n));
    std::atomic<long long> next{0};
    auto work = [&]() {
        for (long long i; (i = next.fetch_add(1)) < n;) f(i);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
}

int writer_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
}

// Closes HDF5 handles in reverse order of creation.
struct Handles {
    std::vector<std::pair<herr_t (*)(hid_t), hid_t>> open;
    hid_t add(herr_t (*close)(hid_t), hid_t id)
    {
        if (id >= 0) open.emplace_back(close, id);
        return id;
    }
    ~Handles()
    {
        for (auto it = open.rbegin(); it != open.rend(); ++it) it->first(it->second);
    }
};

} // namespace

namespace cb {

bool h5_available(std::string &why) { return load_h5(why); }

cb_status h5_write_images(const SnapMeta &m, const SnapGeom &g, const unsigned char *images, std::string &err)
{
    if (!load_h5(err)) return CB_UNSUPPORTED;
    const H5Api &h = g_h5;
    const auto fail = [&](const std::string &what) {
        err = "could not write the snapshot " + m.path + ": " + what;
        return CB_INVALID_ARG;
    };
    if (m.nfields != 2 && m.nfields != kSnapFieldsMax) return fail("nfields must be 2 or 5");
    if (m.deflate < 0 || m.deflate > 9) return fail("deflate level must be 0..9");
    Handles hs;
    const hid_t file = hs.add(h.Fclose, h.Fcreate(m.path.c_str(), kAccTrunc, kDefault, kDefault));
    if (file < 0) return fail("H5Fcreate failed");
    hsize_t dims[3], chunk[3];
    for (int d = 0; d < g.D; ++d) {
        dims[d] = (hsize_t)g.n;
        chunk[d] = (hsize_t)g.ch[d];
    }
    const hid_t space = hs.add(h.Sclose, h.Screate_simple(g.D, dims, nullptr));
    const hid_t plist = hs.add(h.Pclose, h.Pcreate(*h.dataset_create));
    if (space < 0 || plist < 0) return fail("dataspace or property list");
    if (h.Pset_chunk(plist, g.D, chunk) < 0 || h.Pset_shuffle(plist) < 0 ||
        h.Pset_deflate(plist, (unsigned)m.deflate) < 0)
        return fail("chunk / shuffle / deflate properties");

    const size_t cbytes = g.chunk_bytes();
    std::vector<std::vector<unsigned char>> packed((size_t)g.nchunks);
    std::vector<int> zst((size_t)g.nchunks);
    for (int f = 0; f < m.nfields; ++f) {
        const hid_t ds = h.Dcreate2(file, kSnapNames[f], *h.native_float, space, kDefault, plist, kDefault);
        if (ds < 0) return fail(std::string("H5Dcreate2 ") + kSnapNames[f]);
        Handles dsh;
        dsh.add(h.Dclose, ds);
        const unsigned char *img = images + (size_t)f * g.dataset_bytes();
        parallel_for(g.nchunks, writer_threads(), [&](long long q) {
            std::vector<unsigned char> &z = packed[(size_t)q];
            uLongf len = compressBound((uLong)cbytes);
            z.resize(len);
            zst[(size_t)q] = compress2(z.data(), &len, img + (size_t)q * cbytes, (uLong)cbytes, m.deflate);
            z.resize(len);
        });
        for (long long q = 0; q < g.nchunks; ++q) {
            if (zst[(size_t)q] != Z_OK) return fail("zlib compress2 failed");
            // chunk q's first element, row-major over the chunk grid
            hsize_t off[3];
            long long rem = q;
            for (int d = g.D - 1; d >= 0; --d) {
                off[d] = (hsize_t)((rem % g.nch[d]) * g.ch[d]);
                rem /= g.nch[d];
            }
            const std::vector<unsigned char> &z = packed[(size_t)q];
            if (h.Dwrite_chunk(ds, kDefault, 0u, off, z.size(), z.data()) < 0)
                return fail(std::string("H5Dwrite_chunk ") + kSnapNames[f]);
        }
    }
    const hid_t scalar = hs.add(h.Sclose, h.Screate(kScalar));
    if (scalar < 0) return fail("scalar dataspace");
    const std::pair<const char *, double> attrs[] = {{"D", (double)m.D},   {"L", m.L},
                                                     {"nGrids", (double)m.nGrids}, {"reduc", (double)g.reduc},
                                                     {"c", m.c},           {"t", m.t},
                                                     {"a", m.a}};
    for (const auto &at : attrs) {
        const hid_t a = h.Acreate2(file, at.first, *h.native_double, scalar, kDefault, kDefault);
        if (a < 0) return fail(std::string("attribute ") + at.first);
        const herr_t w = h.Awrite(a, *h.native_double, &at.second);
        h.Aclose(a);
        if (w < 0) return fail(std::string("attribute ") + at.first);
    }
    return CB_OK;
}

} // namespace cb

extern "C" {

cb_status cb_hdf5_available(void)
{
    std::string why;
    if (load_h5(why)) return CB_OK;
    set_error(why);
    return CB_UNSUPPORTED;
}

cb_status cb_write_snapshot_h5(const char *path, int D, int nGrids, double L, double c, double t, double a,
                               int nfields, const float *const *fields, int deflate)
{
    if (!path || !fields || D < 1 || D > 3 || nGrids < 2 || (nfields != 2 && nfields != cb::kSnapFieldsMax)) {
        set_error("invalid argument (D 1..3, nGrids >= 2, nfields 2 or 5)");
        return CB_INVALID_ARG;
    }
    for (int f = 0; f < nfields; ++f)
        if (!fields[f]) { set_error("a field pointer is NULL"); return CB_INVALID_ARG; }
    const cb::SnapGeom g = cb::snap_geom(D, nGrids);
    // shuffle the points into the chunk images (the device pack's layout)
    std::vector<unsigned char> img((size_t)nfields * g.dataset_bytes(), 0);
    const long long n = g.n, ce = g.chunk_elems;
    parallel_for((g.points + n - 1) / n, writer_threads(), [&](long long row) {
        for (long long o = row * n; o < (row + 1) * n; ++o) {
            long long cc[3] = {0, 0, 0}, rem = o;
            for (int d = D - 1; d >= 0; --d) {
                cc[d] = rem % n;
                rem /= n;
            }
            long long q = 0, e = 0;
            for (int d = 0; d < D; ++d) {
                q = q * g.nch[d] + cc[d] / g.ch[d];
                e = e * g.ch[d] + cc[d] % g.ch[d];
            }
            for (int f = 0; f < nfields; ++f) {
                uint32_t bits;
                std::memcpy(&bits, fields[f] + o, 4);
                unsigned char *p =
                    img.data() + (size_t)f * g.dataset_bytes() + (size_t)q * g.chunk_bytes() + (size_t)e;
                for (int b = 0; b < 4; ++b) p[(size_t)b * ce] = (unsigned char)(bits >> (8 * b));
            }
        }
    });
    cb::SnapMeta m{path, D, nGrids, L, c, t, a, nfields, deflate};
    std::string err;
    const cb_status r = cb::h5_write_images(m, g, img.data(), err);
    if (r != CB_OK) set_error(err);
    return r;
}

} // extern "C"
