// cb_snapshot.hpp -- HDF5 field snapshots (internal): the saved-point
// geometry of Bubble::OutField (src/bubble.cpp:887-954), its chunk layout,
// the device pack kernel's launch and the file writer.
#pragma once

#include <string>

#include <hip/hip_runtime.h>

#include "../../include/coolbubble.h"
#include "cb_common.hpp"

namespace cb {

// Saved points and chunks of one snapshot dataset.  Large lattices are thinned
// to every reduc-th point per axis (2D above 2048^2, 3D above 512^3); n points
// per axis are kept (rounded up).  Chunks are whole rows: 1D the whole field,
// 2D min(n, max(1, 65536 / n)) rows, 3D one plane (src/bubble.cpp:893-925).
// A chunk's image is its chunk_elems floats after HDF5's shuffle filter:
// byte b of element e at b * chunk_elems + e; a dataset is nchunks images in
// row-major chunk order; rows past the last one of a partial chunk are zero.
struct SnapGeom {
    int D, nGrids, reduc, n;
    long long ch[3];            // chunk extent per axis
    long long nch[3];           // chunks per axis
    long long chunk_elems, nchunks, points;
    __host__ __device__ size_t chunk_bytes() const { return (size_t)chunk_elems * 4; }
    __host__ __device__ size_t dataset_bytes() const { return (size_t)nchunks * chunk_bytes(); }
    __host__ __device__ bool padded() const { return nchunks * chunk_elems != points; }
};

SnapGeom snap_geom(int D, int nGrids);

// Dataset names in file order: phi, pi, then (memorysave = 0) the energy maps.
constexpr int kSnapFieldsMax = 5;
extern const char *const kSnapNames[kSnapFieldsMax];

struct SnapMeta {
    std::string path;
    int D, nGrids;
    double L, c, t, a;
    int nfields;                // 2 or 5
    int deflate;                // gzip level 0..9 (the reference uses 9)
};

// Loads the HDF5 C library on first use; false (with the reason) when absent.
bool h5_available(std::string &why);

// Writes one snapshot file from nfields shuffled dataset images laid out one
// after another (SnapGeom).  Chunks are deflated on host threads and stored
// with H5Dwrite_chunk; the filter pipeline is shuffle + deflate as declared.
// Safe to call from any thread: a failure's reason goes to err.
cb_status h5_write_images(const SnapMeta &m, const SnapGeom &g, const unsigned char *images, std::string &err);

// Device pack: the saved points of phi, pi and (nfields = 5) the local
// energy densities of measureEnergies (src/bubble.cpp:790-833), as float in
// the shuffled chunk images.  The energy maps read phi_m, pi_m and the
// factors a dx and a_kin^(-2D) of msc (the state the last energy
// measurement saw).
hipError_t launch_snapshot_pack(const SimConst &k, int fp32, const SnapGeom &g, int nfields, const void *phi,
                                const void *pi, const void *phi_m, const void *pi_m, const DevScalars *msc,
                                unsigned char *out, hipStream_t st);

// Wall tracking's line read-out (src/bubble.cpp:587-591): out[i] = phi at
// lattice point (c0 + i mod N, c1, c2), i in [0, N) (1D: phi[i]), as double.
hipError_t launch_gather_line(const SimConst &k, int fp32, const void *phi, int c0, int c1, int c2, double *out,
                              hipStream_t st);

} // namespace cb
