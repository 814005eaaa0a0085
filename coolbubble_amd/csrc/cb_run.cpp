// cb_run.cpp -- the reference's run driver over the C ABI: the time loop of
// src/main.cpp:229-310 with the reference's text outputs in a run folder.
//
// Files (same names, layouts and printf formats as the reference, so its
// readers and plotting scripts take them unchanged):
//   parameters.dat       OutParameters, src/bubble.cpp:1018-1032, plus the
//                        "Elapsed time" line of src/main.cpp:303-308
//   EnergyDensities.dat  InitOutputFiles header (src/bubble.cpp:958-966) and
//                        one outEP row (:977-989) at t = 0 and every
//                        OutEnergyEvery steps
//   __xns.dat            mode 1: the nucleation points as GenerateNucleationPoints
//                        writes them (:397-417); mode 0 reads it (:420-453)
//   _phi_k_t=NNNN.dat    FourierTransform + ComputePowerSpectrum (:642-709) at
//                        the initial state and every snapshot
//   FullField.dat        1D: outFullField (:993-1006) at t = 0 and every
//                        Full1DFieldEvery steps
//   _phi_t=NNNN.h5       OutField (:883-954) with every spectrum: the device
//                        packs the saved points into shuffled chunk images, a
//                        side stream copies them to pinned host memory and a
//                        host thread deflates and writes the file while the
//                        time steps go on
//   gamma.dat            wall tracking (TrackWall, :535-559; outGamma,
//                        :1009-1014): a device gather of one lattice line of
//                        phi, scanned on the host as WallFinder does (:576-621)
// Not written: plots, videos.  The library prints nothing (the reference's
// progress lines go to stdout).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <thread>
#include <vector>

#include "cb_sim.hpp"
#include "cb_snapshot.hpp"

using cb::set_error;

namespace {

// std::to_string(num) cut to `prec` decimals (to_precistr, src/main.cpp:63-66)
std::string to_precistr(double num, int prec)
{
    const std::string s = std::to_string(num);
    return s.substr(0, s.find('.') + prec + 1);
}

std::string folder_name(const cb_run_opts &o)
{
    const cb_params &p = o.params;
    std::string n = "Sim" + std::to_string(p.D) + "_D_" + "nt_" + std::to_string(o.nt) + "_dt_" + to_precistr(p.dt, 5) +
                    "_nGrids_" + std::to_string(p.nGrids) + "_L_" + to_precistr(p.L, 1) + "_Mode_" +
                    std::to_string(p.readxns) + "_nB_" + std::to_string(p.nBubbles) + "_c_" + to_precistr(p.c, 2) +
                    "_FFT_" + std::to_string(o.fft);
    if (p.D == 1) n += "_Full1DFieldEvery_" + std::to_string(o.full1d_every);
    n += "_Mpl_" + std::to_string((int)p.Mpl);
    return n;
}

// SnapshotSchedule, src/main.cpp:201-223
std::vector<int> snapshot_schedule(const cb_run_opts &o)
{
    const double dt = o.params.dt, L = o.params.L;
    const int n = o.snapshots, nt = o.nt;
    const double T = nt * dt;
    std::vector<double> times;
    if (o.snapshot_spacing == 1 && T > L && n >= 2) {
        const int nEarly = (n + 1) / 2, nLate = n - nEarly;
        for (int k = 1; k <= nEarly; k++) times.push_back(L * k / nEarly);
        for (int k = 1; k <= nLate; k++) times.push_back(L + (T - L) * k / nLate);
    } else {
        for (int k = 1; k <= n; k++) times.push_back(T * k / n);
    }
    std::vector<int> steps;
    for (double time : times) {
        int st = (int)std::lround(time / dt);
        st = st < 1 ? 1 : st;
        st = st > nt ? nt : st;
        if (steps.empty() || st > steps.back()) steps.push_back(st);
    }
    return steps;
}

struct File {
    FILE *f = nullptr;
    File(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

cb_status open_fail(const std::string &path)
{
    set_error("could not open " + path + " for writing");
    return CB_INVALID_ARG;
}

cb_status write_energy_header(const std::string &folder, double c)
{
    const std::string path = folder + "/EnergyDensities.dat";
    File fp(path, "w");
    if (!fp.f) return open_fail(path);
    std::fprintf(fp.f, "# c and DeltaV are: %.2f %.4f \n", c, 1.0);
    std::fprintf(fp.f, "# columns: 1 t, 2 a, 3 Etot, 4 Ekin, 5 Egrad, 6 Epot, 7-10 spread of Etot/Ekin/Egrad/Epot,\n");
    std::fprintf(fp.f, "#          11 <phi>, 12 spread of <phi>, 13 <phi'>, 14 spread of <phi'>, 15 H, 16 Hubble "
                       "constraint C = H^2 - rho/(3 Mpl^2)\n");
    return CB_OK;
}

cb_status append_energy_row(const std::string &folder, cb_sim *s)
{
    cb_energies e;
    cb_status r = cb_energies_row(s, &e);
    if (r != CB_OK) return r;
    const std::string path = folder + "/EnergyDensities.dat";
    File fp(path, "a");
    if (!fp.f) return open_fail(path);
    const double *v = e.col;
    std::fprintf(fp.f, "%.6e %.10e %.12e %.12e %.12e %.12e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.8e %.8e\n", v[0],
                 v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11], v[12], v[13], v[14], v[15]);
    return CB_OK;
}

cb_status write_parameters(const std::string &folder, const cb_run_opts &o, const cb_sim *s)
{
    const cb_params &p = o.params;
    const std::string path = folder + "/parameters.dat";
    File fp(path, "w");
    if (!fp.f) return open_fail(path);
    const double dx = (double)p.L / p.nGrids;
    const double kUV = std::sqrt(double(p.D)) * M_PI / dx;                  // src/bubble.hpp:54
    const double kIR = 2 * M_PI / double(p.L);                              // :55
    const double kRStar = std::sqrt(double(p.D)) * 4 * M_PI / (double(p.L - 2 * s->Rn));   // src/bubble.cpp:166
    const double kMeff = s->mT;                                             // :170
    std::fprintf(fp.f, "# The data in this folder correspond to the following:\n");
    std::fprintf(fp.f, " D = %d , L = %.10g,  nGrids = %d , TIME_STEP = %.10g  \n", p.D, p.L, p.nGrids, p.dt);
    std::fprintf(fp.f, " c = %.10g , nt = %d,  readxns = %d , expansion = %d  \n", p.c, o.nt, p.readxns, p.expansion);
    std::fprintf(fp.f, " nb = %d , FFT = %d,  ALPHAVAL = %d , WallTracking = %d  \n", p.nBubbles, o.fft, p.alpha,
                 o.wall_tracking);
    std::fprintf(fp.f, "Mpl = %.10g, memorysave = %d, TEST = %d  \n", p.Mpl, o.memorysave, p.test_wave);
    std::fprintf(fp.f, " Full1DFieldEvery = %d , Trad = %.10g , fixedseed = %d \n", o.full1d_every, p.Trad, 1);
    std::fprintf(fp.f, " kUV = %.10g , kIR = %.10g , kRStar = %.10g , kMeff = %.10g \n", kUV, kIR, kRStar, kMeff);
    return CB_OK;
}

cb_status write_spectrum(const std::string &folder, cb_sim *s, int n)
{
    const int cap = (int)(std::sqrt((double)s->p.D) * s->p.nGrids / 2) + 2;
    std::vector<double> k((size_t)cap), kp((size_t)cap), d((size_t)cap);
    int nb = 0;
    cb_status r = cb_power_spectrum(s, k.data(), kp.data(), d.data(), cap, &nb);
    if (r != CB_OK) return r;
    cb_state st;
    r = cb_get_state(s, &st);
    if (r != CB_OK) return r;
    char name[64];
    std::snprintf(name, sizeof name, "/_phi_k_t=%04d.dat", n);
    const std::string path = folder + name;
    File fp(path, "w");
    if (!fp.f) return open_fail(path);
    // std::ostream's default floating format is "%g" (precision 6)
    std::fprintf(fp.f, "# t    a        (then one line per |k| bin: k_comoving  k_physical  Delta(k))\n");
    std::fprintf(fp.f, "%g %g\n", st.t, st.a);
    for (int b = 0; b < nb && b < cap; ++b) std::fprintf(fp.f, "%g   %g   %g\n", k[(size_t)b], kp[(size_t)b], d[(size_t)b]);
    return CB_OK;
}

cb_status append_full_field(const std::string &folder, cb_sim *s)
{
    std::vector<double> phi((size_t)s->nsites), pi((size_t)s->nsites);
    cb_status r = cb_get_fields(s, phi.data(), pi.data());
    if (r != CB_OK) return r;
    const std::string path = folder + "/FullField.dat";
    File fp(path, "a");
    if (!fp.f) return open_fail(path);
    for (double v : phi) std::fprintf(fp.f, "%.5f ", v);
    std::fprintf(fp.f, "\n");
    return CB_OK;
}

// Mode 0 centres from <folder>/__xns.dat: fscanf("%lf %lf %lf") until the end
// or a malformed line (NucleateBubblesAtGivenPoints, src/bubble.cpp:420-453).
cb_status read_xns(const std::string &folder, std::vector<double> &xyz)
{
    const std::string path = folder + "/__xns.dat";
    File fp(path, "r");
    if (!fp.f) {
        set_error("mode 0 needs the bubble positions in " + path + " (one \"x y z\" per line)");
        return CB_INVALID_ARG;
    }
    double v[3];
    while (std::fscanf(fp.f, "%lf %lf %lf", &v[0], &v[1], &v[2]) == 3) xyz.insert(xyz.end(), v, v + 3);
    return CB_OK;
}

#define RUN_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return CB_HIP_ERROR;                                                         \
        }                                                                                \
    } while (0)

// The HDF5 field snapshots of one run.  The energy maps of a snapshot after
// step n are those of the last energy measurement, which saw phi at step
// n - 1 and pi after that step's kick (src/bubble.cpp:761-858), so the run
// keeps phi and the device scalars of step n - 1 (save_map_state) before its
// last step to a snapshot; the initial snapshot's maps are the ctor's
// (src/bubble.cpp:239-265): phi at t = 0, pi before the back-step, a = 1.
struct Snapshots {
    cb_sim *s = nullptr;
    cb::SnapGeom g{};
    int nfields = 2, deflate = 9;
    std::string folder;
    unsigned char *dimg = nullptr;      // device chunk images
    unsigned char *himg = nullptr;      // pinned host copy
    size_t bytes = 0;
    void *phim = nullptr;               // phi of the last energy measurement
    cb::DevScalars *msc = nullptr;      // its factors (a dx, a_kin^(-2D))
    hipStream_t cst = nullptr;          // copy stream
    hipEvent_t packed = nullptr, copied = nullptr;
    std::thread writer;
    cb_status wst = CB_OK;
    std::string werr;

    bool maps() const { return nfields > 2; }

    cb_status init(cb_sim *sim, const cb_run_opts &o, const std::string &dir)
    {
        s = sim;
        folder = dir;
        g = cb::snap_geom(o.params.D, o.params.nGrids);
        nfields = o.memorysave == 0 ? cb::kSnapFieldsMax : 2;
        deflate = o.deflate;
        bytes = (size_t)nfields * g.dataset_bytes();
        RUN_HIP(hipMalloc((void **)&dimg, bytes));
        RUN_HIP(hipHostMalloc((void **)&himg, bytes, hipHostMallocDefault));
        if (maps()) {
            RUN_HIP(hipMalloc(&phim, s->esz * (size_t)s->nsites));
            RUN_HIP(hipMalloc((void **)&msc, sizeof(cb::DevScalars)));
        }
        RUN_HIP(hipStreamCreateWithFlags(&cst, hipStreamNonBlocking));
        RUN_HIP(hipEventCreateWithFlags(&packed, hipEventDisableTiming));
        RUN_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        return CB_OK;
    }

    // Called with the run at step n - 1 before its last step to snapshot n.
    cb_status save_map_state()
    {
        if (!maps()) return CB_OK;
        cb_status r = cb::materialise_drift(s);
        if (r != CB_OK) return r;
        RUN_HIP(hipMemcpyAsync(phim, s->phi[s->iphi], s->esz * (size_t)s->nsites, hipMemcpyDeviceToDevice, s->st));
        RUN_HIP(hipMemcpyAsync(msc, s->dsc, sizeof(cb::DevScalars), hipMemcpyDeviceToDevice, s->st));
        return CB_OK;
    }

    // Waits for the file being written; its status.
    cb_status wait()
    {
        if (writer.joinable()) writer.join();
        if (wst != CB_OK) set_error(werr);
        return wst;
    }

    // Snapshot n of the current state.
    cb_status write(int n)
    {
        cb_status r = wait();   // the images and the pinned buffer are free again
        if (r != CB_OK) return r;
        r = cb::materialise_drift(s);
        if (r != CB_OK) return r;
        const void *phi = s->phi[s->iphi], *pi = s->pi[s->ipi];
        const void *phi_m = phim, *pi_m = pi;
        cb::DevScalars h0{};
        if (maps() && n == 0) {
            if (!s->pi0) { set_error("initial pi was not kept"); return CB_INVALID_ARG; }
            h0.adx = 1.0 * s->k.dx;     // a = 1 in the ctor
            h0.pk2 = 1.0;
            RUN_HIP(hipMemcpyAsync(msc, &h0, sizeof h0, hipMemcpyHostToDevice, s->st));
            phi_m = phi;
            pi_m = s->pi0;
        }
        RUN_HIP(cb::launch_snapshot_pack(s->k, s->fp32, g, nfields, phi, pi, phi_m, pi_m, msc, dimg, s->st));
        RUN_HIP(hipEventRecord(packed, s->st));
        RUN_HIP(hipStreamWaitEvent(cst, packed, 0));
        RUN_HIP(hipMemcpyAsync(himg, dimg, bytes, hipMemcpyDeviceToHost, cst));
        RUN_HIP(hipEventRecord(copied, cst));
        cb::DevScalars h;
        r = cb::sim_scalars(s, h);
        if (r != CB_OK) return r;
        if (n == 0 && s->pi0) {
            hipFree(s->pi0);
            s->pi0 = nullptr;
        }
        char name[32];
        std::snprintf(name, sizeof name, "/_phi_t=%04d.h5", n);
        cb::SnapMeta m{folder + name, s->p.D, s->p.nGrids, s->p.L, s->p.c, s->t, h.a, nfields, deflate};
        writer = std::thread([this, m]() {
            const hipError_t e = hipEventSynchronize(copied);
            if (e != hipSuccess) {
                werr = std::string("snapshot copy: ") + hipGetErrorString(e);
                wst = CB_HIP_ERROR;
                return;
            }
            wst = cb::h5_write_images(m, g, himg, werr);
        });
        return CB_OK;
    }

    ~Snapshots()
    {
        if (writer.joinable()) writer.join();
        if (cst) hipStreamSynchronize(cst);
        if (copied) hipEventDestroy(copied);
        if (packed) hipEventDestroy(packed);
        if (cst) hipStreamDestroy(cst);
        if (dimg) hipFree(dimg);
        if (himg) hipHostFree(himg);
        if (phim) hipFree(phim);
        if (msc) hipFree(msc);
    }
};

// Wall tracking (src/bubble.cpp:525-621).  The field is read along the first
// lattice axis from the tracked bubble's centre; Rmin is the outermost radius
// where it is still above its exit value phiZero, Rmax the first where it is
// at or below the barrier top phiMax; the Lorentz factor is the initial wall
// width over the current one.
struct Wall {
    bool on = false;
    int D = 1, N = 0, it_start = 0, every = 400;
    int centre[3] = {0, 0, 0};
    double phi_zero = 0, phi_max = 0, dx = 0, L = 0, lw = 0, Rn = 0, max_radius = 0;
    double dini = 0, width = 0, rmin = 0, rmax = 0;
    std::vector<double> line;
    double *dline = nullptr;
    std::string path;

    // CanTrackWall (:533-536), the centre (:310-313), WallTrackStart (:562-565)
    cb_status init(const cb_sim *s, const cb_run_opts &o, const std::string &folder)
    {
        const cb_params &p = o.params;
        on = p.D == 1 || p.readxns == -1 || p.readxns == -2;
        if (!on) return CB_OK;
        D = p.D;
        N = p.nGrids;
        every = o.wall_track_every;
        const int mid = D == 3 ? N / 2 : 0;
        if (p.readxns == -2 && D >= 2) centre[0] = centre[1] = N / 2, centre[2] = mid;
        else if (p.readxns == -1 && D >= 2) centre[0] = centre[1] = N / 4, centre[2] = mid;
        const double c = p.c;
        phi_max = c / (6 + 2 * c);
        phi_zero = (2 + c - std::sqrt(4 + c)) / (3 + c);
        dx = s->k.dx;
        L = p.L;
        lw = s->lw;
        Rn = s->Rn;
        max_radius = (p.readxns == -1 && D >= 2) ? L / 4.0 - p.DeltaR : L / 2.0;
        it_start = (int)(M_PI / s->mT / p.dt);
        path = folder + "/gamma.dat";
        line.assign((size_t)N, 0.);
        RUN_HIP(hipMalloc((void **)&dline, sizeof(double) * (size_t)N));
        return CB_OK;
    }

    double gamma(double d) const { return d == 0 ? 0 : dini / d; }

    // WallFinder (:576-621)
    cb_status find(cb_sim *s)
    {
        if (rmax >= 0.98 * max_radius || rmin >= 0.98 * max_radius) return CB_OK;   // left the box
        cb_status r = cb::materialise_drift(s);
        if (r != CB_OK) return r;
        RUN_HIP(cb::launch_gather_line(s->k, s->fp32, s->phi[s->iphi], centre[0], centre[1], centre[2], dline,
                                       s->st));
        RUN_HIP(hipMemcpyAsync(line.data(), dline, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, s->st));
        RUN_HIP(hipStreamSynchronize(s->st));
        // centres lie on cell corners, field values at cell centres
        const auto radius = [&](int i) { return dx * std::sqrt((i + 0.5) * (i + 0.5) + 0.25 * (D - 1)); };
        const int istart = std::max((int)((rmin - lw) / dx), (int)(Rn / (2 * dx)));
        rmin = 0;
        rmax = 0;
        for (int i = istart; (rmax == 0 || rmin == 0) && radius(i) < max_radius; i++) {
            const double v = line[(size_t)(i % N)];
            if (v > phi_zero) rmin = radius(i);
            if (v <= phi_max && rmax == 0) rmax = radius(i);
        }
        if (rmax < rmin) return CB_OK;                     // no wall found: the width stays
        if (rmax * rmin == 0 && rmin + rmax != 0) rmax = rmin;
        width = rmax - rmin;
        return CB_OK;
    }

    cb_status out(double t) const
    {
        File fp(path, "a");
        if (!fp.f) return open_fail(path);
        std::fprintf(fp.f, "%.3f %.4f %.3f %.3f %.3f\n ", t, gamma(width), width, rmin, rmax);
        return CB_OK;
    }

    // TrackWall (:543-559), after step it
    cb_status track(cb_sim *s, int it)
    {
        if (!on) return CB_OK;
        cb_status r = CB_OK;
        if (it == it_start) {
            r = find(s);
            dini = width;
            if (r == CB_OK) r = out(s->t);
        } else if (it > it_start && it % every == 0 && s->t < L / 2) {
            r = find(s);
            if (r == CB_OK) r = out(s->t);
        }
        return r;
    }

    // UpdateWall (:569-573), before a snapshot
    cb_status update(cb_sim *s, int it)
    {
        if (on && it > it_start && s->t < L / 2) return find(s);
        return CB_OK;
    }

    // next step after it at which track() acts, or INT_MAX
    int next_event(int it) const
    {
        if (!on) return INT_MAX;
        if (it < it_start) return it_start;
        return (it / every + 1) * every;
    }

    ~Wall()
    {
        if (dline) hipFree(dline);
    }
};

struct SimGuard {
    cb_sim *s = nullptr;
    ~SimGuard()
    {
        if (s) cb_destroy(s);
    }
};

} // namespace

extern "C" {

cb_status cb_default_run_opts(cb_run_opts *o)
{
    if (!o) { set_error("invalid argument"); return CB_INVALID_ARG; }
    std::memset(o, 0, sizeof *o);
    cb_default_params(&o->params);
    o->nt = 1000;            // nt_default, src/def.h:32
    o->snapshots = 20;       // nSnapshots, :66
    o->snapshot_spacing = 0; // SnapshotSpacing, :67
    o->energy_every = 40;    // OutEnergyEvery, :65
    o->fft = 1;              // FFT, :70
    o->full1d_every = 100;   // Full1DFieldEvery, :71
    o->field_snapshots = 1;  // OutField at every snapshot (src/main.cpp:192-198)
    o->memorysave = 0;       // memorysave, :64
    o->deflate = 9;          // setDeflate(9), src/bubble.cpp:928
    o->wall_tracking = 0;    // WallTracking, :72
    o->wall_track_every = 400; // WallTrackEvery, :73
    o->out_dir = "./";       // Path_default, :63
    return CB_OK;
}

cb_status cb_run_folder_name(const cb_run_opts *o, char *buf, int cap)
{
    if (!o || !buf || cap < 1) { set_error("invalid argument"); return CB_INVALID_ARG; }
    const std::string n = folder_name(*o);
    if ((int)n.size() >= cap) { set_error("buffer too small for the run folder name"); return CB_INVALID_ARG; }
    std::memcpy(buf, n.c_str(), n.size() + 1);
    return CB_OK;
}

cb_status cb_run(cb_run_opts *o)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!o || !o->out_dir) { set_error("invalid argument"); return CB_INVALID_ARG; }
    if (o->nt < 1 || o->snapshots < 1 || o->energy_every < 1 || o->full1d_every < 1 ||
        (o->fft != 0 && o->fft != 1) || (o->snapshot_spacing != 0 && o->snapshot_spacing != 1) ||
        (o->field_snapshots != 0 && o->field_snapshots != 1) || (o->memorysave != 0 && o->memorysave != 1) ||
        o->deflate < 0 || o->deflate > 9 || (o->wall_tracking != 0 && o->wall_tracking != 1) ||
        o->wall_track_every < 1) {
        set_error("nt, snapshots, energy_every, full1d_every and wall_track_every must be >= 1; fft, "
                  "snapshot_spacing, field_snapshots, memorysave and wall_tracking 0 or 1; deflate 0..9");
        return CB_INVALID_ARG;
    }
    if (o->field_snapshots) {
        std::string why;
        if (!cb::h5_available(why)) {
            set_error("field snapshots: " + why);
            return CB_UNSUPPORTED;
        }
    }
    std::string why;
    if (!cb::params_ok(&o->params, why)) { set_error(why); return CB_INVALID_ARG; }
    cb_status r = cb_check_stability(&o->params);
    if (r != CB_OK) return r;
    const std::string folder = (std::filesystem::path(o->out_dir) / folder_name(*o)).string();
    if (folder.size() >= sizeof o->folder) { set_error("run folder path too long"); return CB_INVALID_ARG; }
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    if (ec) { set_error("could not create " + folder + ": " + ec.message()); return CB_INVALID_ARG; }
    std::memcpy(o->folder, folder.c_str(), folder.size() + 1);

    std::vector<double> xyz;
    const double *xns = nullptr;
    int n_xns = 0;
    if (o->params.readxns == 0 && o->params.D >= 2) {
        if (o->xns) {
            xns = o->xns;
            n_xns = o->n_xns;
        } else {
            r = read_xns(folder, xyz);
            if (r != CB_OK) return r;
            xns = xyz.data();
            n_xns = (int)(xyz.size() / 3);
        }
    }
    // Bubble s(folder): the ctor writes __xns.dat (mode 1) and the file headers
    SimGuard g;
    const bool maps = o->field_snapshots && o->memorysave == 0;
    r = cb::create_sim(&o->params, xns, n_xns, maps, &g.s);
    if (r != CB_OK) return r;
    cb_sim *s = g.s;
    Snapshots snap;   // destroyed before the simulation: its writer reads nothing of it
    if (o->field_snapshots) {
        r = snap.init(s, *o, folder);
        if (r != CB_OK) return r;
    }
    Wall wall;
    if (o->wall_tracking) {
        // InitOutputFiles (:967-973): the header, also when the mode cannot be tracked
        const std::string path = folder + "/gamma.dat";
        File fp(path, "w");
        if (!fp.f) return open_fail(path);
        std::fputs("#t, gamma factor, wall width, rmin, rmax\n", fp.f);
        r = wall.init(s, *o, folder);
        if (r != CB_OK) return r;
    }
    if (o->params.readxns == 1 && o->params.D >= 2 && !o->params.test_wave) {
        const std::string path = folder + "/__xns.dat";
        File fp(path, "w");
        if (!fp.f) return open_fail(path);
        std::fputs(s->xns_text.c_str(), fp.f);
    }
    r = write_energy_header(folder, o->params.c);
    if (r == CB_OK) r = write_parameters(folder, *o, s);
    if (r == CB_OK) r = append_energy_row(folder, s);
    if (r == CB_OK && o->field_snapshots) r = snap.write(0);
    if (r == CB_OK && o->fft) r = write_spectrum(folder, s, 0);
    if (r == CB_OK && o->params.D == 1) r = append_full_field(folder, s);
    if (r != CB_OK) return r;

    const std::vector<int> snaps = snapshot_schedule(*o);
    size_t next = 0;
    int it = 0;
    while (it < o->nt) {
        // next event: an energy row, a snapshot, a 1D field line, the end
        int ev = o->nt;
        const int e40 = (it / o->energy_every + 1) * o->energy_every;
        if (e40 < ev) ev = e40;
        if (next < snaps.size() && snaps[next] < ev) ev = snaps[next];
        ev = std::min(ev, wall.next_event(it));
        if (o->params.D == 1) {
            const int ef = (it / o->full1d_every + 1) * o->full1d_every;
            if (ef < ev) ev = ef;
        }
        if (o->field_snapshots && next < snaps.size() && ev == snaps[next]) {
            r = cb_step(s, ev - 1 - it);
            if (r == CB_OK) r = snap.save_map_state();
            if (r == CB_OK) r = cb_step(s, 1);
        } else {
            r = cb_step(s, ev - it);
        }
        if (r != CB_OK) return r;
        it = ev;
        r = wall.track(s, it);
        if (r == CB_OK && o->params.D == 1 && it % o->full1d_every == 0) r = append_full_field(folder, s);
        if (r == CB_OK && it % o->energy_every == 0) r = append_energy_row(folder, s);
        if (r == CB_OK && next < snaps.size() && it == snaps[next]) {
            ++next;
            r = wall.update(s, it);
            if (r == CB_OK && o->field_snapshots) r = snap.write((int)next);
            if (r == CB_OK && o->fft) r = write_spectrum(folder, s, (int)next);
        }
        if (r != CB_OK) return r;
    }
    if (o->field_snapshots) {
        r = snap.wait();
        if (r != CB_OK) return r;
    }
    o->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const std::string path = folder + "/parameters.dat";
    File fp(path, "a");
    if (!fp.f) return open_fail(path);
    std::fprintf(fp.f, "Elapsed time %f seconds\n", o->elapsed_s);
    return CB_OK;
}

} // extern "C"
