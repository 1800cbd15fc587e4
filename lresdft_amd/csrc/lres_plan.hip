// Host side of liblres_hip.so: per-geometry plan, window tables, the C-ABI
// entry points of include/lres.h, inertia bisection and the Brent mu solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/lres.h"

#include "lres_internal.h"

using lres::WinDesc;

namespace {

thread_local std::string g_err;

struct Fail {
    std::string msg;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail{std::string(what) + ": " + hipGetErrorString(e)};
}

void require(bool ok, const std::string& msg) {
    if (!ok) throw Fail{msg};
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hip_check(hipMalloc((void**)&p, count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t st) {
        if (count > n) alloc(count);
        if (count) hip_check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st), "upload");
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

constexpr int kMaxRegisterN = 128;    // register-resident Gauss-Jordan up to n = 128

// Kernel shape of one bucket.  wide 0 / 1: vector-ALU Gauss-Jordan, T x T tile
// per thread on a G x G thread grid, register budget 128 / 512; wide 2: matrix
// cores, T = 16 x 16 tiles per dimension, G = waves per (window, pole group).
struct Shape {
    int T = 0, G = 0, wide = 0;
};
constexpr int kMfma = 2;

// Which pole kernel family runs: the matrix-core kernel in both precisions for
// windows of up to 128 functions (in fp64 its four-pivot blocks need a quarter
// of the vector kernel's barriers: 1.71 against 2.08 ms per pass at 1024
// electrons, although the fp64 MFMA rate only equals the fp64 vector rate),
// the vector-ALU kernels beyond.  LRES_POLE_KERNEL=mfma or =valu selects one
// family for both precisions.  One family per bucket and precision: the
// autotuner only chooses among variants that give bitwise identical results.
bool use_mfma(int /*fp32*/) {
    const char* e = getenv("LRES_POLE_KERNEL");
    if (e && std::string(e) == "valu") return false;
    return true;
}

// The window-fused complex64 form (G = 0) is one of the autotuner's candidates where it
// exists (the tuner then times the per-pole-slot candidates together with k_reduce of the
// bucket's windows, which the fused form does inside its kernel); LRES_POLE_FUSED=1 makes
// it the only candidate, =0 leaves it out.  Both forms give the same bits.
int fused_mode() {
    const char* e = getenv("LRES_POLE_FUSED");
    if (!e) return 2;
    return std::string(e) == "1" ? 1 : std::string(e) == "0" ? 0 : 2;
}

std::vector<Shape> mfma_shapes(int nmax, int fp32) {
    std::vector<Shape> out;
    const int nt = (nmax + 15) / 16;
    if (nt < 2 || nt > (fp32 ? 16 : 8)) return out;
    // LRES_MFMA_WAVES=w restricts the candidates to w waves per (window, pole
    // group), spills allowed (used to check that every variant gives the same bits)
    const char* e = getenv("LRES_MFMA_WAVES");
    const int force = e ? atoi(e) : 0;
    // variants with at most a few spilled registers (128 bytes per lane: values
    // kept per pole, such as the pole weights and base pointers, which the
    // autotuner's timing weighs like everything else); if every variant spills
    // more, the one that spills least, provided that is at most 256 bytes per lane
    int least = -1, least_bytes = 1 << 30;
    for (int nwc = 1; nwc <= 8; ++nwc) {
        if (nwc > nt || (nt > 8 && nwc != 8) || (force > 0 && nwc != force)) continue;
        const int scratch = mfma_scratch_bytes(fp32, nt, nwc);
        if (scratch < 0) continue;
        if (scratch <= 128 || force > 0) out.push_back({nt, nwc, kMfma});
        if (scratch < least_bytes) least = nwc, least_bytes = scratch;
    }
    if (out.empty() && least > 0 && least_bytes <= 256) out.push_back({nt, least, kMfma});
    // complex64: the window-fused form (G = 0: four waves per window, one pole each, the
    // reduction in the kernel; the same bits as one wave per matrix with k_reduce).  Its
    // spills sit outside the block sweep (values kept across a pole), so up to 256 bytes
    // of them are allowed
    if (fp32 && nt <= 8 && force <= 0 && fused_mode() != 0) {
        const int scratch = mfma_scratch_bytes(1, nt, 0);
        if (scratch >= 0 && scratch <= 256) {
            if (fused_mode() == 1) out.clear();
            out.push_back({nt, 0, kMfma});
        }
    }
    return out;
}

// Valid pole-kernel shapes for windows of at most nmax functions.
std::vector<Shape> pole_shapes(int nmax, int fp32) {
    if (use_mfma(fp32)) {
        auto m = mfma_shapes(nmax, fp32);
        if (!m.empty()) return m;
    }
    std::vector<Shape> out;
    for (int T = 2; T <= 6; ++T) {
        const int G = (nmax + T - 1) / T;
        if (G < 2 || G > 32) continue;
        out.push_back({T, G, 0});
        if (G * G <= 256) out.push_back({T, G, 1});
    }
    return out;
}

Shape default_pole_shape(int nmax, int fp32) {
    const auto v = pole_shapes(nmax, fp32);
    if (fp32 && fused_mode() == 1)
        for (const auto& s : v)
            if (s.wide == kMfma && s.G == 0) return s;
    for (const auto& s : v)
        if ((s.wide == kMfma && s.G == 2) || (s.wide == 0 && s.T == 4)) return s;
    return v.empty() ? Shape{} : v.front();
}

hipError_t launch_pole(const Shape& sh, int fp32, const WinDesc* wins, const int32_t* ids, int nids, const double* SW,
                       const double* HW, int M, int ngroup, const double* circ, double* Dpart, double* Fpart,
                       hipStream_t st, int acc = 0, lres::WinOut out = {}) {
    if (sh.wide == kMfma)
        return launch_pole_mfma(fp32, sh.T, sh.G, wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, acc, out,
                                st);
    return launch_pole_gj(sh.T, sh.G, sh.wide, fp32, wins, ids, nids, SW, HW, M, ngroup, circ, Dpart, Fpart, st);
}

Shape inertia_shape(int nmax) {
    for (int T = 2; T <= 8; ++T) {
        const int G = (nmax + T - 1) / T;
        if (G <= 32) return {T, std::max(G, 2)};
    }
    return {};
}

}  // namespace

struct lres_plan {
    int64_t nb = 0, ng = 0, nwin = 0;
    double dV = 0.0;
    int M = 0;
    int precision = LRES_FP64;
    int ngroup = 1;
    int kmax = 0;
    hipStream_t st = nullptr;
    std::vector<WinDesc> wins;
    std::vector<int32_t> n_of, k_of;
    struct Bucket {
        int nmax = 0;
        std::vector<int32_t> ids;
        int64_t id_off = 0;
        Shape shape[2];        // pole kernel shape per precision (0 = fp64, 1 = fp32)
        bool tuned[2] = {false, false};
        Shape ishape;          // inertia kernel shape
        bool big = false;      // n above the register-resident limit: blocked sweep in a workspace
        int bmax = 0;          // panel width of the blocked sweep
        int kmax = 0;          // largest |T_B| of the bucket
    };
    std::vector<Bucket> buckets;
    bool acc_lds = true;       // fp64 group partials accumulated in LDS (LRES_POLE_ACC != "global")
    bool autotune = true;
    bool tune_log = false;
    DevBuf<WinDesc> d_wins;
    DevBuf<int64_t> d_Wflat, d_Woff;
    DevBuf<int32_t> d_pos;                      // local window index -> T-last position
    DevBuf<int32_t> d_bucket_ids;
    DevBuf<int64_t> d_ws_off;                   // workspace offset per window (big windows), cx<double> units
    DevBuf<double> d_work;                      // workspace of the blocked sweep (2 doubles per entry)
    DevBuf<double> d_SW, d_HW, d_Q, d_D, d_F, d_Dpart, d_Fpart, d_NB, d_EB, d_sums;
    // complex64 matrix-core passes: one float pair per pole and upper-triangle entry
    // ((M / 2) k x k pairs per window), and which windows' partials are in that form
    DevBuf<float> d_slots;
    DevBuf<int8_t> d_by_pole;
    std::vector<int8_t> by_pole;
    DevBuf<int64_t> d_own_g, d_xptr;
    int dens_blocks6 = 0;                        // owned rows of X are whole 6-function atom blocks
    // the same rows for the density kernel in 64-point chunks, block-interleaved: per
    // window its first chunk (d_chk_off), per chunk its first block slot (d_chk_bs), per
    // block slot 64 lanes of the block's T-local column (d_chk_idx) and 6 x 64 values
    // (d_chk_val); rows shorter than the chunk's longest are padded with zero blocks
    DevBuf<int64_t> d_chk_off, d_chk_bs;
    DevBuf<int32_t> d_chk_idx;
    DevBuf<double> d_chk_val;
    DevBuf<int32_t> d_xidx;
    DevBuf<double> d_xval;
    DevBuf<int64_t> d_Hp, d_Hi;
    DevBuf<double> d_Hv;
    DevBuf<unsigned long long> d_itot;
    DevBuf<int32_t> d_cand, d_icnt;              // lambda_min: candidate windows, per-window counts
    DevBuf<double> d_rq;                         // lambda_min: per-window Rayleigh bound
    // lambda_min on the device: search state, shifts, totals, candidate list and its
    // compaction buffer
    DevBuf<LminState> d_lmin;
    DevBuf<double> d_lsig;
    DevBuf<unsigned long long> d_litot;
    DevBuf<int32_t> d_lcand, d_ltmp;
    std::vector<int32_t> bucket_of;              // window -> bucket index
    std::vector<int64_t> fn_at;                  // per window (Woff offsets): global function at each T-last position
    DevBuf<double> d_adjG, d_adjZ;               // adjoint workspaces (complex entries as double pairs)
    DevBuf<int32_t> d_adj_ids;
    DevBuf<int64_t> d_adj_off;
    // H(v) assembly on the device (lres_plan_set_basis / lres_set_potential)
    bool have_basis = false;
    int nf_max = 0;
    int64_t npair = 0;
    DevBuf<int32_t> d_pairA, d_pairB, d_ent_r, d_ent_a, d_ent_b, d_nf_of, d_hmap;
    DevBuf<int64_t> d_pent_off, d_xat_off, d_vp_off;
    DevBuf<double> d_xat, d_vpair, d_TW;
    bool have_H = false;
    // keep-edge bound (k_edge_*): the potential of the current H (vcur, copied by every
    // lres_set_potential), the potential of the last lambda_min (vref) with its lowest
    // eigenvalue bracket in d_edge; pot_h: the current H was assembled from a potential;
    // edge_ok: d_edge and vref describe a lambda_min on such an H
    DevBuf<double> d_vcur, d_vref, d_edge_part, d_edge;
    DevBuf<int> d_edge_skip;
    bool pot_h = false;
    // forces (lres_force_setup / lres_band_forces): the basis on per-atom boxes, the
    // atom pairs whose boxes meet with their window entries, grid-point owners
    bool have_forces = false;
    int f_nf = 0, f_nbox = 0;
    int64_t f_natom = 0, f_npair = 0;
    int64_t f_shape[3] = {0, 0, 0};
    DevBuf<int64_t> d_forigin, d_fent_ptr, d_fent_off, d_faptr, d_faent;
    DevBuf<int32_t> d_fpA, d_fpB, d_fent_n, d_fowner, d_frow, d_ftpos;
    DevBuf<double> d_fpos, d_fexps, d_fL, d_fX, d_fDX, d_fTX, d_fGH, d_fGS, d_fpair, d_fq;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // side streams (default priority) over which independent per-bucket launches of a
    // pass or an inertia count are spread, with their fork and join events; created on
    // first use on the plan's device
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSide] = {};
    double timing[3] = {0, 0, 0};
    // cumulative pole-kernel statistics per precision: passes and milliseconds (HIP events)
    int64_t stat_passes[2] = {0, 0};
    double stat_ms[2] = {0.0, 0.0};
    // pinned, mapped, coherent host staging (d_pin: its device address): a pass's
    // sums are written there by the sum kernel itself (slots 0, 1), inertia totals
    // come back in slots 2.. (an asynchronous copy), and the inertia kernels read
    // their shifts from the block at kPinSlots; every call that uses it waits for its
    // stream before returning
    static constexpr int kPinSlots = 2 + 1024;
    // after the two blocks above: the device step's state (StepState), published there
    // by its last controller stage
    static constexpr int kStepPin = 64;
    // after that: the parameter block of the next device step (StepPrm order), written by
    // the host before the step is launched or its graph replayed
    static constexpr int kPrmPin = 16;
    static constexpr int kPrmOff = 2 * kPinSlots + kStepPin;
    double* h_pin = nullptr;
    double* d_pin = nullptr;
    // the circle of the pass in flight in device memory: {c, R, skip, skip_density}
    // (written by k_set_circle for host-decided passes, by the mu controller in a
    // device step)
    DevBuf<double> d_circ;
    // queued warm density step (lres_step_enqueue / lres_step_finish): its state, the
    // trial slots per batch and a pair of timing events per slot
    DevBuf<lres::StepState> d_step;
    bool step_queued = false;
    int step_depth = 0;
    bool step_timed = false;                     // the queued step's slots carry timing events
    DevBuf<double> d_ssig;                        // the keep-edge shifts of a device step
    std::vector<hipEvent_t> ev_slot;
    ~lres_plan() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_join)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_slot)
            if (e) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (auto& q : side)
            if (q) (void)hipStreamDestroy(q);
        if (h_pin) (void)hipHostFree(h_pin);
    }
};

namespace {

// Independent launches spread over the plan's side streams: side_fork makes every
// side stream wait for the work queued so far on the plan stream, side_stream(P, i) is
// the stream of the i-th launch, side_join makes the plan stream wait for the first
// `used` side streams.  Each launch writes outputs of its own windows only.
void side_fork(lres_plan* P) {
    if (!P->ev_fork) {
        hip_check(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming), "event");
        for (int i = 0; i < lres_plan::kSide; ++i) {
            hip_check(hipStreamCreateWithFlags(&P->side[i], hipStreamNonBlocking), "stream");
            hip_check(hipEventCreateWithFlags(&P->ev_join[i], hipEventDisableTiming), "event");
        }
    }
    hip_check(hipEventRecord(P->ev_fork, P->st), "event");
    for (int i = 0; i < lres_plan::kSide; ++i) hip_check(hipStreamWaitEvent(P->side[i], P->ev_fork, 0), "wait");
}
hipStream_t side_stream(const lres_plan* P, int i) { return P->side[i % lres_plan::kSide]; }
void side_join(lres_plan* P, int used) {
    for (int i = 0; i < std::min(used, lres_plan::kSide); ++i) {
        hip_check(hipEventRecord(P->ev_join[i], P->side[i]), "event");
        hip_check(hipStreamWaitEvent(P->st, P->ev_join[i], 0), "wait");
    }
}

int fail_from(const Fail& f) {
    g_err = f.msg;
    return 1;
}

template <typename F>
int guarded(F&& fn) {
    try {
        g_err.clear();
        fn();
        return 0;
    } catch (const Fail& f) {
        return fail_from(f);
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// The keep-edge bound replaces the two inertia counts of a warm step when it proves
// their outcome (k_step_edge); LRES_EDGE_BOUND=0 switches it off (the counts always
// run; same decisions, so the same results).  kEdgeSafety (Ha) keeps a proved decision
// clear of the rounding of the factorisations it replaces.
constexpr double kEdgeSafety = 1e-8;

bool edge_bound_enabled() {
    const char* e = getenv("LRES_EDGE_BOUND");
    return !(e && std::string(e) == "0");
}

// no keep-edge reference (the current H is not H(v) of the plan's basis): a NaN
// bracket, which proves no decision
void edge_clear(lres_plan* P) {
    if (P->d_edge.n >= kEdgeState) hip_check(launch_edge_ref(P->d_edge.p, NAN, NAN, P->st), "keep-edge bound");
}

void gather_h(lres_plan* P, const int64_t* Hp, const int64_t* Hi, const double* Hv) {
    hip_check(hipEventRecord(P->ev[0], P->st), "event");
    hip_check(launch_gather_csr(P->d_wins.p, (int)P->nwin, P->d_Wflat.p, P->d_Woff.p, P->d_pos.p, Hp, Hi, Hv,
                                P->d_HW.p, P->st),
              "gather kernel");
    hip_check(hipEventRecord(P->ev[1], P->st), "event");
    P->have_H = true;
    P->pot_h = false;
    edge_clear(P);
}

// Windows above the register-resident limit of the vector kernels: in fp32 the
// matrix-core kernel holds up to 256 functions (16 x 16 tiles over eight waves);
// otherwise (fp64, or larger) the blocked sweep in a workspace.
bool big_on_mfma(int nmax, int fp32) {
    return fp32 && nmax <= 256 && use_mfma(1) && !mfma_shapes(nmax, 1).empty();
}

// Density kernel input: the owned rows of X, with their chunked form (one lane per point
// over atom blocks) when the rows allow it, unless LRES_DENSITY_KERNEL=rows (the
// row-per-wave kernel, any row shape)
DensRows density_rows(const lres_plan* P) {
    const char* e = std::getenv("LRES_DENSITY_KERNEL");
    const bool chunks = P->dens_blocks6 && !(e && std::strcmp(e, "rows") == 0);
    return DensRows{P->d_own_g.p, P->d_xptr.p, P->d_xidx.p, P->d_xval.p, chunks ? P->d_chk_off.p : nullptr,
                    P->d_chk_bs.p, P->d_chk_idx.p, P->d_chk_val.p};
}

// The partial-sum buffer a pole kernel of shape sh writes: the per-pole slots for
// the complex64 matrix-core kernel (allocated on first use), else the group sums.
double* pole_partials(lres_plan* P, const Shape& sh, int fp32) {
    if (!(fp32 && sh.wide == kMfma && sh.G > 0)) return P->d_Dpart.p;
    const size_t need = P->d_Q.n * (size_t)(P->M / 2) * 2;
    if (P->d_slots.n < need) P->d_slots.alloc(need);
    return reinterpret_cast<double*>(P->d_slots.p);
}

// fp64 matrix-core launches with more than one pole per group accumulate the group's
// D and F partial sums in LDS (packed upper triangles, 16 B per entry) and store them
// once, when that fits beside a second workgroup on the CU (at most 64 KB: |T_B| <= 89);
// the same additions in the same order as the read-modify-write of global memory, so
// the same bits.  LRES_POLE_ACC=global keeps the read-modify-write.
int pole_acc_entries(const lres_plan* P, const lres_plan::Bucket& b, const Shape& sh, int fp32) {
    if (fp32 && sh.wide == kMfma && sh.G == 0) return b.kmax * (b.kmax + 1) / 2;   // window-fused: window sums
    if (fp32 || sh.wide != kMfma || P->M / 2 <= P->ngroup || !P->acc_lds) return 0;
    const int u = b.kmax * (b.kmax + 1) / 2;
    return (size_t)u * 16 <= 64 * 1024 ? u : 0;
}

// Autotuner: time every valid shape of a bucket on the loaded H (kTuneLaunches launches
// after one warm-up, HIP events) and keep the fastest.  All shapes apply the
// same operations to every matrix entry, so the choice never changes results.
// When the window-fused complex64 form is a candidate, the per-pole-slot candidates
// are timed with k_reduce of the bucket's windows (the others marked as reduced),
// the work the fused form does inside its kernel.
// four launches: enough to order candidates a few percent apart (the complex64 slot form with
// its k_reduce and the window-fused form come within about 2% of each other)
constexpr int kTuneLaunches = 4;

void tune_bucket(lres_plan* P, lres_plan::Bucket& b, int fp32, double c, double R) {
    const auto cands = pole_shapes(b.nmax, fp32);
    require(!cands.empty(), "no kernel shape for windows of " + std::to_string(b.nmax) + " functions");
    b.shape[fp32] = default_pole_shape(b.nmax, fp32);
    b.tuned[fp32] = true;
    if (!P->autotune || cands.size() == 1) return;
    auto fused = [&](const Shape& sh) { return fp32 && sh.wide == kMfma && sh.G == 0; };
    bool with_reduce = false;
    for (const auto& sh : cands) with_reduce = with_reduce || fused(sh);
    std::vector<int8_t> mark;
    if (with_reduce) {
        mark.assign(P->nwin, 2);
        for (int32_t w : b.ids) mark[w] = 1;
        P->d_by_pole.upload(mark.data(), mark.size(), P->st);
        P->by_pole.clear();   // prepare_passes uploads the passes' table after tuning
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    float best = 1e30f;
    hip_check(launch_set_circle(P->d_circ.p, c, R, 0.0, 1.0, P->st), "circle");
    for (const auto& sh : cands) {
        auto launch = [&] {
            hip_check(launch_pole(sh, fp32, P->d_wins.p, P->d_bucket_ids.p + b.id_off, (int)b.ids.size(), P->d_SW.p,
                                  P->d_HW.p, P->M, P->ngroup, P->d_circ.p, pole_partials(P, sh, fp32), P->d_Fpart.p,
                                  P->st, pole_acc_entries(P, b, sh, fp32),
                                  lres::WinOut{P->d_D.p, P->d_F.p, P->d_Q.p, P->d_NB.p, P->d_EB.p}),
                      "pole kernel (tuning)");
            if (with_reduce && !fused(sh))
                hip_check(launch_reduce(P->d_wins.p, (int)P->nwin, P->ngroup, P->d_Dpart.p, P->d_Fpart.p,
                                        P->d_slots.p, P->M / 2, P->d_by_pole.p, P->d_Q.p, P->d_D.p, P->d_F.p,
                                        P->d_NB.p, P->d_EB.p, P->d_circ.p, P->st),
                          "reduce kernel (tuning)");
        };
        launch();
        hip_check(hipEventRecord(e0, P->st), "event");
        for (int i = 0; i < kTuneLaunches; ++i) launch();
        hip_check(hipEventRecord(e1, P->st), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        if (P->tune_log)
            fprintf(stderr, "[lres tune] nmax=%d windows=%zu %s %s T=%d G=%d: %.4f ms/launch%s\n", b.nmax,
                    b.ids.size(), fp32 ? "fp32" : "fp64",
                    sh.wide == kMfma ? (sh.G == 0 ? "mfma-fused" : "mfma") : (sh.wide ? "valu512" : "valu128"), sh.T,
                    sh.G, ms / kTuneLaunches, with_reduce && !fused(sh) ? " (with k_reduce)" : "");
        if (ms < best) {
            best = ms;
            b.shape[fp32] = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// The buckets whose pole solves run on the blocked sweep, as one span of the size-ordered
// id list: first = index of the first such bucket (-1 when they are not a tail of the
// bucket list, or there are none), with the span's ids, largest n and narrowest panel.
struct BigSpan {
    int first = -1;
    int64_t id_off = 0;
    int nids = 0, nmax = 0, bmax = 0;
};
BigSpan big_span(const lres_plan* P, int fp32) {
    BigSpan sp;
    const int nb = (int)P->buckets.size();
    int first = nb;
    while (first > 0) {
        const auto& b = P->buckets[first - 1];
        if (!(b.big && !big_on_mfma(b.nmax, fp32))) break;
        --first;
    }
    if (first == nb) return sp;
    for (int i = 0; i < first; ++i) {
        const auto& b = P->buckets[i];
        if (b.big && !big_on_mfma(b.nmax, fp32)) return sp;   // not a tail: per-bucket launches
    }
    sp.first = first;
    sp.id_off = P->buckets[first].id_off;
    sp.bmax = P->buckets[first].bmax;
    for (int i = first; i < nb; ++i) {
        const auto& b = P->buckets[i];
        sp.nids += (int)b.ids.size();
        sp.nmax = std::max(sp.nmax, b.nmax);
        sp.bmax = std::min(sp.bmax, b.bmax);
    }
    return sp;
}

// Before the passes of one precision: every register-kernel bucket tuned (the
// autotuner times its variants at the circle (c, R) on the plan stream) and the
// reduction's table of windows whose partials are per-pole slots (complex64
// matrix-core buckets) on the device; both change only on a bucket's first pass.
void prepare_passes(lres_plan* P, int fp32, double c, double R) {
    std::vector<int8_t> by_pole(P->nwin, 0);
    for (auto& b : P->buckets) {
        if (b.big && !big_on_mfma(b.nmax, fp32)) continue;
        if (!b.tuned[fp32]) tune_bucket(P, b, fp32, c, R);
        if (fp32 && b.shape[fp32].wide == kMfma)
            for (int32_t w : b.ids) by_pole[w] = b.shape[fp32].G == 0 ? 2 : 1;
    }
    // complex64 passes only (fp64 partials are always group sums)
    if (fp32 && by_pole != P->by_pole) {
        P->d_by_pole.upload(by_pole.data(), by_pole.size(), P->st);
        hip_check(hipStreamSynchronize(P->st), "sync");
        P->by_pole = by_pole;
    }
}

// A complex64 pass whose windows are all reduced inside the window-fused pole kernel
// (by_pole 2 throughout the prepared table) launches no k_reduce.
bool reduce_needed(const lres_plan* P, int fp32) {
    if (!fp32 || P->by_pole.size() != (size_t)P->nwin) return true;
    for (int8_t v : P->by_pole)
        if (v != 2) return true;
    return false;
}

// Queue one pass at the circle in P->d_circ: pole solves, reduction, the density into
// d_dens if given (skipped on the device when d_circ[3] != 0) and the sums of N_B and
// E_B into `sums` (device or mapped host memory).  Every kernel leaves at once when
// d_circ[2] != 0.  The buckets must be prepared (prepare_passes).  e0 / e1 (optional):
// events recorded on the plan stream around the pole kernels.  step_stage >= 0: a slot
// of a device step, whose sum kernel then runs that controller stage (k_sum_step).
void queue_pass(lres_plan* P, double* d_dens, double* d_NB, double* d_EB, double* sums, hipEvent_t e0 = nullptr,
                hipEvent_t e1 = nullptr, int step_stage = -1) {
    const int fp32 = P->precision == LRES_MIXED ? 1 : 0;
    const double* circ = P->d_circ.p;
    // the buckets on the blocked sweep (size-generic: every window's n comes from its
    // descriptor) go in one launch when they are the tail of the size-ordered id list, so
    // that a spread of window sizes still fills the chip
    const BigSpan big = big_span(P, fp32);
    // the register-kernel buckets; with more than one launch in the pass, they run
    // concurrently on the side streams (each writes its own windows' partial sums)
    // beside the blocked launch
    std::vector<lres_plan::Bucket*> reg;
    int nlaunch = big.first >= 0 ? 1 : 0;
    for (auto& b : P->buckets) {
        if (b.big && !big_on_mfma(b.nmax, fp32)) {
            if (big.first < 0) ++nlaunch;
            continue;
        }
        require(b.tuned[fp32], "pass queued before its buckets were tuned");
        reg.push_back(&b);
        ++nlaunch;
    }
    const bool spread = nlaunch > 1 && !reg.empty();
    double* NB = d_NB ? d_NB : P->d_NB.p;
    double* EB = d_EB ? d_EB : P->d_EB.p;
    if (e0) hip_check(hipEventRecord(e0, P->st), "event");
    if (spread) side_fork(P);
    if (big.first >= 0)
        hip_check(launch_pole_big(fp32, big.bmax, big.nmax, P->d_wins.p, P->d_bucket_ids.p + big.id_off, big.nids,
                                  P->d_SW.p, P->d_HW.p, P->M, P->ngroup, circ, P->d_work.p, P->d_ws_off.p,
                                  P->d_Dpart.p, P->d_Fpart.p, P->st),
                  "blocked pole kernel");
    else
        for (auto& b : P->buckets)
            if (b.big && !big_on_mfma(b.nmax, fp32))
                hip_check(launch_pole_big(fp32, b.bmax, b.nmax, P->d_wins.p, P->d_bucket_ids.p + b.id_off,
                                          (int)b.ids.size(), P->d_SW.p, P->d_HW.p, P->M, P->ngroup, circ,
                                          P->d_work.p, P->d_ws_off.p, P->d_Dpart.p, P->d_Fpart.p, P->st),
                          "blocked pole kernel");
    int used = 0;
    for (auto* bp : reg) {
        auto& b = *bp;
        const hipStream_t sq = spread ? side_stream(P, used++) : P->st;
        hip_check(launch_pole(b.shape[fp32], fp32, P->d_wins.p, P->d_bucket_ids.p + b.id_off, (int)b.ids.size(),
                              P->d_SW.p, P->d_HW.p, P->M, P->ngroup, circ, pole_partials(P, b.shape[fp32], fp32),
                              P->d_Fpart.p, sq, pole_acc_entries(P, b, b.shape[fp32], fp32),
                              lres::WinOut{P->d_D.p, P->d_F.p, P->d_Q.p, NB, EB}),
                  "pole kernel");
    }
    if (spread) side_join(P, used);
    if (e1) hip_check(hipEventRecord(e1, P->st), "event");
    if (reduce_needed(P, fp32))
        hip_check(launch_reduce(P->d_wins.p, (int)P->nwin, P->ngroup, P->d_Dpart.p, P->d_Fpart.p, P->d_slots.p,
                                P->M / 2, fp32 ? P->d_by_pole.p : nullptr, P->d_Q.p, P->d_D.p, P->d_F.p, NB, EB, circ,
                                P->st),
                  "reduce kernel");
    if (d_dens)
        hip_check(launch_density(P->d_wins.p, (int)P->nwin, P->kmax, P->d_D.p, density_rows(P), d_dens, circ + 3,
                                 P->st),
                  "density kernel");
    if (step_stage >= 0)
        hip_check(launch_sum_step(NB, EB, (int)P->nwin, sums, P->d_step.p, step_stage, P->d_circ.p,
                                  P->d_pin + 2 * lres_plan::kPinSlots, P->st),
                  "sum kernel");
    else
        hip_check(launch_sum2(NB, EB, (int)P->nwin, sums, circ + 2, P->st), "sum kernel");
}

// pole solves + reduction (+ density) at (mu, e_lo) decided on the host; sums to host
// if requested
void run_pass(lres_plan* P, double mu, double e_lo, double* d_dens, double* d_NB, double* d_EB, double* h_sums) {
    require(P->have_H, "no Hamiltonian loaded: call lres_set_hamiltonian_csr first");
    require(mu > e_lo, "mu must lie above e_lo");
    require(!P->step_queued, "a queued density step is pending: call lres_step_finish first");
    const double c = 0.5 * (mu + e_lo), R = 0.5 * (mu - e_lo);
    const int fp32 = P->precision == LRES_MIXED ? 1 : 0;
    prepare_passes(P, fp32, c, R);
    hip_check(launch_set_circle(P->d_circ.p, c, R, 0.0, d_dens ? 0.0 : 1.0, P->st), "circle");
    queue_pass(P, d_dens, d_NB, d_EB, h_sums ? P->d_pin : P->d_sums.p, P->ev[2], P->ev[3]);
    if (h_sums) {
        hip_check(hipStreamSynchronize(P->st), "sync");
        h_sums[0] = P->h_pin[0];
        h_sums[1] = P->h_pin[1];
        float t0 = 0, t1 = 0;
        if (hipEventElapsedTime(&t0, P->ev[0], P->ev[1]) == hipSuccess) P->timing[0] = t0;
        if (hipEventElapsedTime(&t1, P->ev[2], P->ev[3]) == hipSuccess) {
            P->timing[1] = t1;
            P->stat_passes[fp32] += 1;
            P->stat_ms[fp32] += t1;
        }
    }
}

// Sylvester inertia counts #{eigenvalues of (H_W, S_W) below sigma} summed over
// windows, at ns shifts; per-window counts into d_counts[w * ns + s] if given.
// cand (optional): only these windows (host list of window indices).
// Inertia counts of windows of 17..128 functions on the matrix cores (block
// elimination, a quarter of the vector kernel's barriers); LRES_INERTIA_KERNEL=valu
// selects the vector-ALU kernel.  Both count the positive pivots of the same
// unpivoted LDL^T.
bool inertia_on_mfma(int nmax) {
    const char* e = getenv("LRES_INERTIA_KERNEL");
    if (e && std::string(e) == "valu") return false;
    return nmax > 16 && nmax <= kMaxRegisterN;
}

// queue the inertia counts at ns shifts (totals into d_itot, per window into d_counts
// if given) without waiting for them
// the count kernels at the ns shifts d_sig (device-readable), totals added into d_itot
// (zeroed by the caller), per-window counts into d_counts if given; skip (may be null):
// every kernel leaves at once when *skip != 0
void inertia_kernels(lres_plan* P, const double* d_sig, int ns, int32_t* d_counts,
                     const std::vector<int32_t>* cand = nullptr, const int* skip = nullptr) {
    std::vector<int32_t> sub;        // candidate ids grouped by bucket
    std::vector<int64_t> sub_off;    // per bucket: first entry in sub
    if (cand) {
        for (const auto& b : P->buckets) {
            sub_off.push_back((int64_t)sub.size());
            for (int32_t w : *cand)
                if (P->bucket_of[w] == (int32_t)(&b - P->buckets.data())) sub.push_back(w);
        }
        sub_off.push_back((int64_t)sub.size());
        if (sub.empty()) sub.push_back(0);   // keeps the upload non-empty; no launch reads it
        P->d_cand.upload(sub.data(), sub.size(), P->st);
    }
    // the big buckets (a tail of the size order) in one launch on the plan stream, as in
    // run_pass; the other buckets concurrently on the side streams when there are several
    // launches (integer totals by atomics, per-window counts disjoint)
    size_t first_big = P->buckets.size();
    while (first_big > 0 && P->buckets[first_big - 1].big) --first_big;
    bool merged = first_big < P->buckets.size();
    for (size_t bi = 0; bi < first_big; ++bi) merged = merged && !P->buckets[bi].big;
    int nother = merged ? 1 : 0;
    for (size_t bi = 0; bi < P->buckets.size(); ++bi) {
        const int nids = cand ? (int)(sub_off[bi + 1] - sub_off[bi]) : (int)P->buckets[bi].ids.size();
        if (nids > 0 && !(merged && P->buckets[bi].big)) ++nother;
    }
    const bool spread = nother > 1;
    if (spread) side_fork(P);
    if (merged) {
        int nmax = 0, bmax = 1 << 30;
        for (size_t bi = first_big; bi < P->buckets.size(); ++bi) {
            nmax = std::max(nmax, P->buckets[bi].nmax);
            bmax = std::min(bmax, P->buckets[bi].bmax);
        }
        const int32_t* ids = cand ? P->d_cand.p + sub_off[first_big] : P->d_bucket_ids.p + P->buckets[first_big].id_off;
        int nids = 0;
        for (size_t bi = first_big; bi < P->buckets.size(); ++bi)
            nids += cand ? (int)(sub_off[bi + 1] - sub_off[bi]) : (int)P->buckets[bi].ids.size();
        if (nids > 0)
            hip_check(launch_inertia_big(bmax, nmax, P->d_wins.p, ids, nids, P->d_SW.p, P->d_HW.p, d_sig, ns,
                                         P->d_work.p, (int64_t)P->d_work.n, d_counts, P->d_itot.p, skip, P->st),
                      "blocked inertia kernel");
    }
    int used = 0;
    for (size_t bi = 0; bi < P->buckets.size(); ++bi) {
        const auto& b = P->buckets[bi];
        const int32_t* ids = cand ? P->d_cand.p + sub_off[bi] : P->d_bucket_ids.p + b.id_off;
        const int nids = cand ? (int)(sub_off[bi + 1] - sub_off[bi]) : (int)b.ids.size();
        if (nids == 0 || (merged && b.big)) continue;
        if (b.big) {   // blocked buckets share the workspace: they stay on the plan stream
            hip_check(launch_inertia_big(b.bmax, b.nmax, P->d_wins.p, ids, nids, P->d_SW.p, P->d_HW.p, d_sig,
                                         ns, P->d_work.p, (int64_t)P->d_work.n, d_counts, P->d_itot.p, skip,
                                         P->st),
                      "blocked inertia kernel");
            continue;
        }
        const hipStream_t sq = spread ? side_stream(P, used++) : P->st;
        if (inertia_on_mfma(b.nmax))
            hip_check(launch_inertia_mfma((b.nmax + 15) / 16, P->d_wins.p, ids, nids, P->d_SW.p, P->d_HW.p,
                                          d_sig, ns, d_counts, P->d_itot.p, skip, sq),
                      "matrix-core inertia kernel");
        else
            hip_check(launch_inertia(b.ishape.T, b.ishape.G, P->d_wins.p, ids, nids, P->d_SW.p, P->d_HW.p,
                                     d_sig, ns, d_counts, P->d_itot.p, skip, sq),
                      "inertia kernel");
    }
    if (spread) side_join(P, used);
}

void inertia_launch(lres_plan* P, const double* sig, int ns, int32_t* d_counts,
                    const std::vector<int32_t>* cand = nullptr) {
    require(P->have_H, "no Hamiltonian loaded");
    require(ns > 0 && ns <= 1024, "nshift must be in 1..1024");
    require(ns <= lres_plan::kPinSlots, "too many shifts");
    std::memcpy(P->h_pin + lres_plan::kPinSlots, sig, ns * sizeof(double));
    hip_check(hipMemsetAsync(P->d_itot.p, 0, ns * sizeof(unsigned long long), P->st), "memset");
    inertia_kernels(P, P->d_pin + lres_plan::kPinSlots, ns, d_counts, cand);
}

void inertia(lres_plan* P, const double* sig, int ns, int64_t* h_total, int32_t* d_counts,
             const std::vector<int32_t>* cand = nullptr, int32_t* h_counts = nullptr) {
    inertia_launch(P, sig, ns, d_counts, cand);
    static_assert(sizeof(unsigned long long) == sizeof(double), "totals share the staging slots");
    const unsigned long long* tot = reinterpret_cast<const unsigned long long*>(P->h_pin + 2);
    hip_check(hipMemcpyAsync(P->h_pin + 2, P->d_itot.p, ns * sizeof(unsigned long long), hipMemcpyDeviceToHost, P->st),
              "copy");
    // per-window counts (counts[w * ns + s]) with the totals, under one synchronisation
    if (h_counts && d_counts)
        hip_check(hipMemcpyAsync(h_counts, d_counts, (size_t)P->nwin * ns * sizeof(int32_t), hipMemcpyDeviceToHost,
                                 P->st),
                  "copy");
    hip_check(hipStreamSynchronize(P->st), "sync");
    for (int s = 0; s < ns; ++s) h_total[s] = (int64_t)tot[s];
}

// Pass slots per batch of a device step (lres_step_enqueue), at most; batches after the
// first hold at least kMoreStepDepth.
constexpr int kMaxStepDepth = 8, kMoreStepDepth = 4;

// One batch of a queued warm density step: `depth` trial slots (a pass at the
// controller's circle, then its feed stage), the final stage (a pass at the accepted mu
// when it was not the last trial; the density either way) and the publish stage, which
// copies the search state to the mapped staging.  The slots the search does not need
// leave at once.  Pole-kernel events per slot for the pass statistics.
void step_batch(lres_plan* P, int depth, double* d_dens, bool timed) {
    if ((int)P->ev_slot.size() < 2 * (kMaxStepDepth + 1)) {
        P->ev_slot.resize(2 * (kMaxStepDepth + 1), nullptr);
        for (auto& e : P->ev_slot)
            if (!e) hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    auto ev = [&](int i) { return timed ? P->ev_slot[i] : nullptr; };
    for (int i = 0; i < depth; ++i)
        queue_pass(P, nullptr, nullptr, nullptr, P->d_sums.p, ev(2 * i), ev(2 * i + 1),
                   i + 1 < depth ? lres::kStepFeed : lres::kStepFeedFinal);
    queue_pass(P, d_dens, nullptr, nullptr, P->d_sums.p, ev(2 * depth), ev(2 * depth + 1), lres::kStepPublish);
}

// is the plan's stream being captured into a graph
bool capturing(lres_plan* P) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(P->st, &cs), "capture status");
    return cs != hipStreamCaptureStatusNone;
}

// the buckets of this precision are tuned (a step can then be captured: tuning runs and
// waits for kernels)
bool passes_prepared(const lres_plan* P, int fp32) {
    for (const auto& b : P->buckets)
        if (!(b.big && !big_on_mfma(b.nmax, fp32)) && !b.tuned[fp32]) return false;
    return true;
}

// The first batch of a device step from the parameter block at P->d_pin + kPrmOff: the
// keep-edge test (lcao3d.py:420; its shifts, zeroed totals and the bound's decision by
// k_step_edge, then the counts unless the bound decided), the controller's start on the
// totals, depth trial slots, the final slot and the publish stage.  No host value of the
// step is baked into the launches, and nothing here allocates, waits or sets memory, so
// the batch can be captured into a graph and replayed with new parameters.
void step_launch(lres_plan* P, int depth, double* d_dens, bool timed) {
    const double* prm = P->d_pin + lres_plan::kPrmOff;
    double* edge = edge_bound_enabled() && P->d_edge.n >= kEdgeState ? P->d_edge.p : nullptr;
    hip_check(launch_step_edge(prm, P->d_ssig.p, P->d_itot.p, edge, kEdgeSafety, P->d_edge_skip.p, P->st),
              "keep-edge test");
    inertia_kernels(P, P->d_ssig.p, 2, nullptr, nullptr, P->d_edge_skip.p);
    hip_check(launch_mu_begin(P->d_step.p, P->d_itot.p, prm, P->d_circ.p, P->st), "mu controller");
    step_batch(P, depth, d_dens, timed);
}

}  // namespace

namespace {

// the dyadic cell width of a lambda_min tolerance: the smallest power of two >= tol
double lmin_width(double tol) {
    int e = 0;
    const double m = std::frexp(tol, &e);   // tol = m 2^e, m in [0.5, 1)
    return m == 0.5 ? tol : std::ldexp(1.0, e);
}

// lambda_min's result from its cell [lo, hi) (no window eigenvalue below lo, one below
// hi): the midpoint; the cell is also the keep-edge reference at the potential of this H
void lambda_min_finish(lres_plan* P, double lo, double hi, double* out) {
    *out = 0.5 * (lo + hi);
    if (P->pot_h) {
        hip_check(hipMemcpyAsync(P->d_vref.p, P->d_vcur.p, P->ng * sizeof(double), hipMemcpyDeviceToDevice, P->st),
                  "copy");
        hip_check(launch_edge_ref(P->d_edge.p, lo, hi, P->st), "keep-edge bound");
    } else {
        edge_clear(P);
    }
}

// lambda_min searched on the device (k_lmin_*) when every window is in one bucket on the
// matrix-core inertia kernel; LRES_LMIN=host keeps the host-driven search (the same
// bracket bit for bit)
bool lmin_on_device(const lres_plan* P) {
    const char* e = getenv("LRES_LMIN");
    if (e && std::string(e) == "host") return false;
    if (P->buckets.size() != 1) return false;
    const auto& b = P->buckets[0];
    return !b.big && inertia_on_mfma(b.nmax) && (int64_t)b.ids.size() == P->nwin;
}

// The device search: probe and bracket, then rounds of (linear count launch, controller)
// queued kLminRounds at a time with one wait per batch.  Returns false (nothing decided)
// when the probes find no count-free shift: the host search then starts over.
constexpr int kLminRounds = 12;

bool lambda_min_device(lres_plan* P, double tol, double* cell_out) {
    const int nwin = (int)P->nwin;
    const auto& b = P->buckets[0];
    const int nt = (b.nmax + 15) / 16;
    if (P->d_lmin.n < 1) P->d_lmin.alloc(1);
    if ((int)P->d_lsig.n < kLminMaxK) P->d_lsig.alloc(kLminMaxK);
    if ((int)P->d_litot.n < kLminMaxK) P->d_litot.alloc(kLminMaxK);
    if ((int)P->d_lcand.n < nwin) P->d_lcand.alloc(nwin);
    if ((int)P->d_ltmp.n < nwin) P->d_ltmp.alloc(nwin);
    if ((int64_t)P->d_icnt.n < (int64_t)nwin * kLminMaxK) P->d_icnt.alloc((size_t)nwin * kLminMaxK);
    if ((int64_t)P->d_rq.n < nwin) P->d_rq.alloc(nwin);
    if (P->d_edge_skip.n < 1) P->d_edge_skip.alloc(1);
    const double width = lmin_width(tol);
    LminState* st = P->d_lmin.p;
    const int32_t* ids = P->d_bucket_ids.p + b.id_off;
    hip_check(launch_diag_ratio(P->d_wins.p, nwin, P->d_SW.p, P->d_HW.p, P->d_rq.p, P->st), "diag kernel");
    hip_check(launch_lmin_probe(P->d_rq.p, nwin, st, tol, P->d_lsig.p, P->d_litot.p, P->st), "lambda_min controller");
    hip_check(launch_inertia_mfma(nt, P->d_wins.p, ids, nwin, P->d_SW.p, P->d_HW.p, P->d_lsig.p, 1 + kLminProbe,
                                  nullptr, P->d_litot.p, nullptr, P->st),
              "matrix-core inertia kernel");
    hip_check(launch_lmin_bracket(st, P->d_lsig.p, P->d_litot.p, ids, nwin, P->d_lcand.p, P->st),
              "lambda_min controller");
    // every round has ncand K <= max(kLminWaves, kLminMinK nwin) (window, shift) pairs
    const int G = std::max(kLminWaves, kLminMinK * nwin);
    const int* lin = &st->ncand;
    for (int batch = 0;; ++batch) {
        for (int r = 0; r < kLminRounds; ++r) {
            hip_check(launch_inertia_mfma(nt, P->d_wins.p, P->d_lcand.p, G, P->d_SW.p, P->d_HW.p, P->d_lsig.p,
                                          kLminMaxK, P->d_icnt.p, P->d_litot.p, nullptr, P->st, lin),
                      "matrix-core inertia kernel");
            hip_check(launch_lmin_round(st, P->d_lsig.p, P->d_litot.p, P->d_icnt.p, kLminMaxK, P->d_lcand.p,
                                        P->d_ltmp.p, P->st),
                      "lambda_min controller");
        }
        // the cell: one count at the grid point inside the final bracket when it has one
        int* skip = P->d_edge_skip.p;
        hip_check(launch_lmin_snap(st, width, P->d_lsig.p, P->d_litot.p, skip, P->st), "lambda_min controller");
        hip_check(launch_inertia_mfma(nt, P->d_wins.p, ids, nwin, P->d_SW.p, P->d_HW.p, P->d_lsig.p, 1, nullptr,
                                      P->d_litot.p, skip, P->st),
                  "matrix-core inertia kernel");
        hip_check(launch_lmin_snap_end(st, width, P->d_lsig.p, P->d_litot.p, skip, P->st), "lambda_min controller");
        LminState h;
        hip_check(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, P->st), "copy");
        hip_check(hipStreamSynchronize(P->st), "sync");
        require(h.fail != 1, "inertia count at the Rayleigh bound is zero (indefinite S?)");
        if (h.fail == 2) return false;
        if (h.done) {
            *cell_out = h.cell;
            return true;
        }
        require(batch < 16, "lambda_min: no convergence of the multisection");
    }
}

}  // namespace

extern "C" {

const char* lres_last_error(void) { return g_err.c_str(); }

int lres_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

#pragma clang fp contract(off)
// lres_build_windows / lres_build_windows_device: the same tables; the Voronoi owners
// of the grid points (ng x ncell distances) on the host or by k_voronoi_owner.
static int build_windows_impl(const double* cells, int64_t ncell, const double* apos, int64_t natom,
                              const int64_t* atom_of_fn, int64_t nb, const double* box, const int64_t* shape,
                              double dx, double R_w, int64_t* Wflat, int64_t* Woff, int64_t* Oflat, int64_t* Ooff,
                              bool on_device) {
    return guarded([&] {
        require(ncell > 0 && natom > 0 && nb > 0, "empty geometry");
        // minimum image exactly as d - L * round(d / L) with round half to even
        auto mi = [&](double d, int ax) { return d - box[ax] * std::nearbyint(d / box[ax]); };
        std::vector<int64_t> cell_of_atom(natom);
        for (int64_t a = 0; a < natom; ++a) {
            double best = std::numeric_limits<double>::infinity();
            int64_t bi = 0;
            for (int64_t B = 0; B < ncell; ++B) {
                double s = 0.0;
                for (int ax = 0; ax < 3; ++ax) {
                    const double d = mi(apos[3 * a + ax] - cells[3 * B + ax], ax);
                    s += d * d;
                }
                const double r = std::sqrt(s);
                if (r < best) {
                    best = r;
                    bi = B;
                }
            }
            cell_of_atom[a] = bi;
        }
        std::vector<int64_t> cell_of_fn(nb);
        for (int64_t f = 0; f < nb; ++f) {
            require(atom_of_fn[f] >= 0 && atom_of_fn[f] < natom, "atom_of_fn out of range");
            cell_of_fn[f] = cell_of_atom[atom_of_fn[f]];
        }
        // W(B): functions whose cell centre lies within R_w of B
        Woff[0] = 0;
        std::vector<char> near(ncell);
        for (int64_t B = 0; B < ncell; ++B) {
            for (int64_t C = 0; C < ncell; ++C) {
                double s = 0.0;
                for (int ax = 0; ax < 3; ++ax) {
                    const double d = mi(cells[3 * C + ax] - cells[3 * B + ax], ax);
                    s += d * d;
                }
                near[C] = std::sqrt(s) <= R_w;
            }
            int64_t cnt = 0;
            for (int64_t f = 0; f < nb; ++f)
                if (near[cell_of_fn[f]]) {
                    if (Wflat) Wflat[Woff[B] + cnt] = f;
                    ++cnt;
                }
            Woff[B + 1] = Woff[B] + cnt;
        }
        if (!Wflat) return;   // sizing call: Woff only
        // Voronoi owners of the grid points, ties to the lowest B (strict '<' in ascending B)
        const int64_t n1 = shape[0], n2 = shape[1], n3 = shape[2], ng = n1 * n2 * n3;
        std::vector<int64_t> owner(ng);
        if (on_device) {
            require(ncell <= 4096, "more than 4096 cells for the device Voronoi kernel");
            int ndev = 0;
            hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
            require(ndev > 0, "no HIP device");
            DevBuf<double> d_cells;
            DevBuf<int32_t> d_owner;
            d_cells.upload(cells, (size_t)3 * ncell, nullptr);
            d_owner.alloc((size_t)ng);
            hip_check(launch_voronoi_owner(d_cells.p, (int)ncell, box, shape, dx, d_owner.p, nullptr), "voronoi kernel");
            std::vector<int32_t> o32(ng);
            hip_check(hipMemcpy(o32.data(), d_owner.p, ng * sizeof(int32_t), hipMemcpyDeviceToHost), "copy");
            for (int64_t g = 0; g < ng; ++g) owner[g] = o32[g];
        }
        for (int64_t g = 0; g < ng && !on_device; ++g) {
            const int64_t i = g / (n2 * n3), j = (g / n3) % n2, k = g % n3;
            const double r[3] = {(double)i * dx, (double)j * dx, (double)k * dx};
            double best = std::numeric_limits<double>::infinity();
            int64_t bi = 0;
            for (int64_t B = 0; B < ncell; ++B) {
                const double d0 = mi(r[0] - cells[3 * B + 0], 0);
                const double d1 = mi(r[1] - cells[3 * B + 1], 1);
                const double d2 = mi(r[2] - cells[3 * B + 2], 2);
                const double s = d0 * d0 + d1 * d1 + d2 * d2;
                if (s < best) {
                    best = s;
                    bi = B;
                }
            }
            owner[g] = bi;
        }
        std::vector<int64_t> cnt(ncell + 1, 0);
        for (int64_t g = 0; g < ng; ++g) cnt[owner[g] + 1]++;
        Ooff[0] = 0;
        for (int64_t B = 0; B < ncell; ++B) Ooff[B + 1] = Ooff[B] + cnt[B + 1];
        std::vector<int64_t> fill(Ooff, Ooff + ncell);
        for (int64_t g = 0; g < ng; ++g) Oflat[fill[owner[g]]++] = g;
    });
}
#pragma clang fp contract(on)

int lres_build_windows(const double* cells, int64_t ncell, const double* apos, int64_t natom,
                       const int64_t* atom_of_fn, int64_t nb, const double* box, const int64_t* shape, double dx,
                       double R_w, int64_t* Wflat, int64_t* Woff, int64_t* Oflat, int64_t* Ooff) {
    return build_windows_impl(cells, ncell, apos, natom, atom_of_fn, nb, box, shape, dx, R_w, Wflat, Woff, Oflat,
                              Ooff, false);
}

int lres_build_windows_device(const double* cells, int64_t ncell, const double* apos, int64_t natom,
                              const int64_t* atom_of_fn, int64_t nb, const double* box, const int64_t* shape,
                              double dx, double R_w, int64_t* Wflat, int64_t* Woff, int64_t* Oflat, int64_t* Ooff) {
    return build_windows_impl(cells, ncell, apos, natom, atom_of_fn, nb, box, shape, dx, R_w, Wflat, Woff, Oflat,
                              Ooff, true);
}

int lres_plan_create(lres_plan** out, int64_t nb, int64_t ng, double dV, int M, int precision,
                     const int64_t* Wflat, const int64_t* Woff, int64_t nwin, const int64_t* Oflat,
                     const int64_t* Ooff, const int64_t* Sp, const int64_t* Si, const double* Sv, const int64_t* Xp,
                     const int64_t* Xi, const double* Xv, void* stream) {
    lres_plan* P = nullptr;
    const int rc = guarded([&] {
        require(out != nullptr, "null plan pointer");
        require(M >= 2 && M % 2 == 0, "M must be even");
        require(nwin > 0 && nb > 0 && ng > 0, "empty plan");
        require(precision == LRES_FP64 || precision == LRES_MIXED, "unknown precision");
        int ndev = 0;
        hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        require(ndev > 0, "no HIP device");
        P = new lres_plan();
        P->nb = nb;
        P->ng = ng;
        P->nwin = nwin;
        P->dV = dV;
        P->M = M;
        P->precision = precision;
        P->st = (hipStream_t)stream;
        for (auto& e : P->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        static_assert(sizeof(lres::StepState) <= lres_plan::kStepPin * sizeof(double), "step state staging");
        static_assert(sizeof(lres::StepState) % 8 == 0, "step state copied as 8-byte words");
        static_assert(kStepPrmCount <= lres_plan::kPrmPin, "step parameter block");
        hip_check(hipHostMalloc((void**)&P->h_pin, (lres_plan::kPrmOff + lres_plan::kPrmPin) * sizeof(double),
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc");
        hip_check(hipHostGetDevicePointer((void**)&P->d_pin, P->h_pin, 0), "hipHostGetDevicePointer");
        P->wins.resize(nwin);
        P->n_of.resize(nwin);
        P->k_of.resize(nwin);
        P->fn_at.assign(Woff[nwin] - Woff[0], -1);
        std::vector<int32_t> pos_all(Woff[nwin]);
        std::vector<int32_t> loc(nb, -1);
        std::vector<double> SW_all, Q_all;
        std::vector<int64_t> own_g, xptr(1, 0);
        std::vector<int32_t> xidx;
        std::vector<double> xval;
        int64_t mat_off = 0, kk_off = 0;
        for (int64_t w = 0; w < nwin; ++w) {
            const int64_t* W = Wflat + Woff[w];
            const int n = (int)(Woff[w + 1] - Woff[w]);
            require(n > 0, "empty window");
            for (int a = 0; a < n; ++a) {
                require(W[a] >= 0 && W[a] < nb, "window function index out of range");
                require(a == 0 || W[a] > W[a - 1], "window lists must be strictly ascending");
                loc[W[a]] = a;
            }
            // T_B: window functions with support on own(B) (src/lrft_cpp.cpp:168-177 touches exactly these)
            std::vector<char> inT(n, 0);
            for (int64_t o = Ooff[w]; o < Ooff[w + 1]; ++o) {
                const int64_t g = Oflat[o];
                require(g >= 0 && g < ng, "owned grid index out of range");
                for (int64_t p = Xp[g]; p < Xp[g + 1]; ++p) {
                    const int a = loc[Xi[p]];
                    if (a >= 0) inT[a] = 1;
                }
            }
            int k = 0;
            for (int a = 0; a < n; ++a) k += inT[a];
            const int m = n - k;
            int32_t* pos = pos_all.data() + Woff[w];
            int nxt_nT = 0, nxt_T = m;
            for (int a = 0; a < n; ++a) pos[a] = inT[a] ? nxt_T++ : nxt_nT++;
            for (int a = 0; a < n; ++a) P->fn_at[Woff[w] - Woff[0] + pos[a]] = W[a];
            // dense S_W in T-last order
            const size_t s0 = SW_all.size();
            SW_all.resize(s0 + (size_t)n * n, 0.0);
            for (int a = 0; a < n; ++a)
                for (int64_t p = Sp[W[a]]; p < Sp[W[a] + 1]; ++p) {
                    const int b = loc[Si[p]];
                    if (b >= 0) SW_all[s0 + (size_t)pos[a] * n + pos[b]] = Sv[p];
                }
            // owned rows of X in T-local columns, and Q_B = sum x x^T dV
            const size_t q0 = Q_all.size();
            Q_all.resize(q0 + (size_t)k * k, 0.0);
            for (int64_t o = Ooff[w]; o < Ooff[w + 1]; ++o) {
                const int64_t g = Oflat[o];
                const size_t e0 = xidx.size();
                for (int64_t p = Xp[g]; p < Xp[g + 1]; ++p) {
                    const int a = loc[Xi[p]];
                    if (a < 0) continue;
                    xidx.push_back(pos[a] - m);
                    xval.push_back(Xv[p]);
                }
                for (size_t e = e0; e < xidx.size(); ++e)
                    for (size_t f = e0; f < xidx.size(); ++f)
                        Q_all[q0 + (size_t)xidx[e] * k + xidx[f]] += xval[e] * xval[f] * dV;
                own_g.push_back(g);
                xptr.push_back((int64_t)xidx.size());
            }
            WinDesc d;
            d.mat_off = mat_off;
            d.kk_off = kk_off;
            d.own_off = Ooff[w] - Ooff[0];
            d.n = n;
            d.k = k;
            d.m = m;
            d.nown = (int32_t)(Ooff[w + 1] - Ooff[w]);
            P->wins[w] = d;
            P->n_of[w] = n;
            P->k_of[w] = k;
            P->kmax = std::max(P->kmax, k);
            mat_off += (int64_t)n * n;
            kk_off += (int64_t)k * k;
            for (int a = 0; a < n; ++a) loc[W[a]] = -1;
        }
        // pole groups: enough (window, group) blocks to fill 256 CUs a few times over
        P->ngroup = (int)std::min<int64_t>(M / 2, std::max<int64_t>(1, (2048 + nwin - 1) / nwin));
        // buckets: windows grouped by size in steps of 12 functions (one H2 of the default basis)
        {
            std::vector<int32_t> order(nwin);
            for (int64_t w = 0; w < nwin; ++w) order[w] = (int32_t)w;
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return P->n_of[a] < P->n_of[b]; });
            for (int32_t w : order) {
                const int key = (P->n_of[w] + 11) / 12;
                if (P->buckets.empty() || (P->buckets.back().nmax + 11) / 12 != key) P->buckets.emplace_back();
                auto& b = P->buckets.back();
                b.ids.push_back(w);
                b.nmax = std::max(b.nmax, P->n_of[w]);
                b.kmax = std::max(b.kmax, P->k_of[w]);
            }
        }
        std::vector<int32_t> ids;
        std::vector<int64_t> ws_off(nwin, 0);
        int64_t ws_total = 0;
        P->bucket_of.assign(nwin, -1);
        for (auto& b : P->buckets) {
            for (int32_t w : b.ids) P->bucket_of[w] = (int32_t)(&b - P->buckets.data());
            b.id_off = (int64_t)ids.size();
            ids.insert(ids.end(), b.ids.begin(), b.ids.end());
            b.big = b.nmax > kMaxRegisterN;
            if (b.big) {
                // panel width: the widest of 16 / 8 / 4 whose two n x b complex panels fit the 160 KB of LDS
                b.bmax = 16;
                while (b.bmax > 4 && big_lds_bytes(b.nmax, b.bmax, 16) > 160 * 1024) b.bmax /= 2;
                require(big_lds_bytes(b.nmax, b.bmax, 16) <= 160 * 1024,
                        "window of " + std::to_string(b.nmax) + " functions exceeds the blocked solver's LDS");
                for (int32_t w : b.ids) {
                    ws_off[w] = ws_total;
                    ws_total += (int64_t)P->ngroup * P->n_of[w] * P->n_of[w];
                }
            } else {
                b.ishape = inertia_shape(b.nmax);
                require(!pole_shapes(b.nmax, 0).empty() && !pole_shapes(b.nmax, 1).empty(),
                        "no kernel shape for windows of " + std::to_string(b.nmax));
            }
        }
        P->d_ws_off.upload(ws_off.data(), nwin, P->st);
        // the blocked inertia kernel cuts the same workspace into real slots of the largest
        // blocked window's n^2 and needs one per blocked window at least (its launch takes
        // every blocked bucket at once)
        int64_t ws_len = ws_total * 2;
        int64_t nbig = 0;
        int nmax_big = 0;
        for (const auto& b : P->buckets)
            if (b.big) {
                nbig += (int64_t)b.ids.size();
                nmax_big = std::max(nmax_big, b.nmax);
            }
        ws_len = std::max<int64_t>(ws_len, nbig * nmax_big * nmax_big);
        if (ws_len) P->d_work.alloc((size_t)ws_len);
        const char* pa = std::getenv("LRES_POLE_ACC");
        P->acc_lds = !(pa && std::string(pa) == "global");
        const char* at = std::getenv("LRES_AUTOTUNE");
        P->autotune = !(at && at[0] == '0');
        const char* tl = std::getenv("LRES_TUNE_LOG");
        P->tune_log = tl && tl[0] == '1';
        hipStream_t st = P->st;
        P->d_wins.upload(P->wins.data(), nwin, st);
        P->d_Wflat.upload(Wflat, Woff[nwin], st);
        std::vector<int64_t> woff_rel(nwin + 1);
        for (int64_t w = 0; w <= nwin; ++w) woff_rel[w] = Woff[w] - Woff[0];
        P->d_Woff.upload(woff_rel.data(), nwin + 1, st);
        P->d_pos.upload(pos_all.data(), pos_all.size(), st);
        P->d_bucket_ids.upload(ids.data(), ids.size(), st);
        P->d_SW.upload(SW_all.data(), SW_all.size(), st);
        P->d_HW.alloc(SW_all.size());
        P->d_Q.upload(Q_all.data(), Q_all.size(), st);
        P->d_D.alloc(std::max<size_t>(1, Q_all.size()));
        P->d_F.alloc(std::max<size_t>(1, Q_all.size()));
        // passes write the upper triangles only; the density kernel's zero padding blocks
        // multiply lower-triangle entries by zero, so those start as zeros (never NaN)
        hip_check(hipMemsetAsync(P->d_D.p, 0, P->d_D.n * sizeof(double), P->st), "memset");
        hip_check(hipMemsetAsync(P->d_F.p, 0, P->d_F.n * sizeof(double), P->st), "memset");
        P->d_Dpart.alloc(std::max<size_t>(1, Q_all.size() * P->ngroup));
        P->d_Fpart.alloc(std::max<size_t>(1, Q_all.size() * P->ngroup));
        P->d_NB.alloc(nwin);
        P->d_EB.alloc(nwin);
        P->d_sums.alloc(2);
        P->d_circ.alloc(4);
        P->d_step.alloc(1);
        P->d_ssig.alloc(2);
        // inertia totals at their largest (1024 shifts): a captured step holds this address
        P->d_itot.alloc(1024);
        if (P->d_edge_skip.n < 1) P->d_edge_skip.alloc(1);
        // whole atom blocks in every owned row: 6 consecutive T-local columns per
        // block, blocks ascending (the density kernel's lane-per-point form)
        P->dens_blocks6 = 1;
        for (size_t o = 0; o + 1 < xptr.size() && P->dens_blocks6; ++o) {
            const int64_t r0 = xptr[o], nnz = xptr[o + 1] - r0;
            if (nnz % 6 != 0) {
                P->dens_blocks6 = 0;
                break;
            }
            for (int64_t q = r0; q < r0 + nnz; q += 6) {
                for (int i = 1; i < 6; ++i)
                    if (xidx[q + i] != xidx[q] + i) P->dens_blocks6 = 0;
                if (q > r0 && xidx[q] < xidx[q - 6] + 6) P->dens_blocks6 = 0;   // ascending blocks
            }
        }
        if (P->dens_blocks6) {
            // the chunked form of the owned rows (lane-contiguous loads in the density
            // kernel): padding blocks repeat the row's last column with zero values, so
            // they add exact zeros after the row's own terms
            std::vector<int64_t> chk_off(nwin + 1, 0), bs(1, 0);
            std::vector<int32_t> cidx;
            std::vector<double> cval;
            for (int64_t w = 0; w < nwin; ++w) {
                const WinDesc& dw = P->wins[w];
                chk_off[w] = (int64_t)bs.size() - 1;
                for (int c0 = 0; c0 < dw.nown; c0 += 64) {
                    const int np = std::min(64, dw.nown - c0);
                    int nbc = 0;
                    for (int l = 0; l < np; ++l) {
                        const int64_t o = dw.own_off + c0 + l;
                        nbc = std::max(nbc, (int)((xptr[o + 1] - xptr[o]) / 6));
                    }
                    const int64_t base = bs.back();
                    cidx.resize((size_t)(base + nbc) * 64, 0);
                    cval.resize((size_t)(base + nbc) * 6 * 64, 0.0);
                    for (int l = 0; l < np; ++l) {
                        const int64_t o = dw.own_off + c0 + l, r0 = xptr[o];
                        const int nb = (int)((xptr[o + 1] - r0) / 6);
                        for (int A = 0; A < nbc; ++A) {
                            const bool real = A < nb;
                            const int64_t q = r0 + 6 * (real ? A : std::max(nb - 1, 0));
                            cidx[(size_t)(base + A) * 64 + l] = nb > 0 ? xidx[q] : 0;
                            for (int i = 0; i < 6; ++i)
                                cval[((size_t)(base + A) * 6 + i) * 64 + l] = real ? xval[q + i] : 0.0;
                        }
                    }
                    bs.push_back(base + nbc);
                }
            }
            chk_off[nwin] = (int64_t)bs.size() - 1;
            if (cidx.empty()) {   // no owned point at all: keep the buffers non-empty
                cidx.push_back(0);
                cval.push_back(0.0);
            }
            P->d_chk_off.upload(chk_off.data(), chk_off.size(), st);
            P->d_chk_bs.upload(bs.data(), bs.size(), st);
            P->d_chk_idx.upload(cidx.data(), cidx.size(), st);
            P->d_chk_val.upload(cval.data(), cval.size(), st);
            hip_check(hipStreamSynchronize(st), "sync");   // the host vectors go out of scope
        }
        P->d_own_g.upload(own_g.data(), own_g.size(), st);
        P->d_xptr.upload(xptr.data(), xptr.size(), st);
        P->d_xidx.upload(xidx.data(), std::max<size_t>(xidx.size(), 0), st);
        P->d_xval.upload(xval.data(), xval.size(), st);
        hip_check(hipStreamSynchronize(st), "sync");
        *out = P;
    });
    if (rc) delete P;
    return rc;
}

int lres_plan_set_basis(lres_plan* P, const int64_t* aof, int64_t natom, const int64_t* Xp, const int64_t* Xi,
                        const double* Xv, const int64_t* Tp, const int64_t* Ti, const double* Tv, const int64_t* Wflat,
                        const int64_t* Woff) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        const int64_t nb = P->nb, ng = P->ng, nwin = P->nwin;
        require(natom > 0, "no atoms");
        std::vector<int32_t> nf(natom, 0), fidx(nb);
        for (int64_t f = 0; f < nb; ++f) {
            require(aof[f] >= 0 && aof[f] < natom, "atom_of_fn out of range");
            fidx[f] = nf[aof[f]]++;
        }
        int nf_max = 0;
        for (auto x : nf) nf_max = std::max(nf_max, (int)x);
        require(nf_max <= 8, "at most 8 basis functions per atom are supported by the H(v) kernel");
        // pass 1: support size per atom (grid points where any of its functions is stored)
        std::vector<int64_t> last(natom, -1), supp(natom, 0);
        for (int64_t r = 0; r < ng; ++r)
            for (int64_t p = Xp[r]; p < Xp[r + 1]; ++p) {
                const int64_t A = aof[Xi[p]];
                if (last[A] != r) {
                    last[A] = r;
                    supp[A]++;
                }
            }
        std::vector<int64_t> xoff(natom + 1, 0);
        for (int64_t A = 0; A < natom; ++A) xoff[A + 1] = xoff[A] + supp[A] * nf[A];
        std::vector<double> xat(xoff[natom], 0.0);
        // pass 2: atom-blocked values and the common-support lists of atom pairs
        std::fill(last.begin(), last.end(), -1);
        std::vector<int64_t> cur(natom, -1);   // support index of r for atom A (valid when last[A] == r)
        std::vector<int64_t> fillc(natom, 0);
        std::vector<int32_t> at_r;
        struct PairList {
            std::vector<int32_t> r, a, b;
        };
        std::vector<PairList> lists;
        std::vector<std::pair<int32_t, int32_t>> pair_atoms;
        std::unordered_map<int64_t, int32_t> pid;
        for (int64_t r = 0; r < ng; ++r) {
            at_r.clear();
            for (int64_t p = Xp[r]; p < Xp[r + 1]; ++p) {
                const int64_t f = Xi[p];
                const int64_t A = aof[f];
                if (last[A] != r) {
                    last[A] = r;
                    cur[A] = fillc[A]++;
                    at_r.push_back((int32_t)A);
                }
                xat[xoff[A] + cur[A] * nf[A] + fidx[f]] = Xv[p];
            }
            std::sort(at_r.begin(), at_r.end());
            for (size_t i = 0; i < at_r.size(); ++i)
                for (size_t j = i; j < at_r.size(); ++j) {
                    const int32_t A = at_r[i], B = at_r[j];
                    const int64_t key = (int64_t)A * natom + B;
                    auto it = pid.find(key);
                    int32_t id;
                    if (it == pid.end()) {
                        id = (int32_t)lists.size();
                        pid.emplace(key, id);
                        lists.emplace_back();
                        pair_atoms.emplace_back(A, B);
                    } else {
                        id = it->second;
                    }
                    lists[id].r.push_back((int32_t)r);
                    lists[id].a.push_back((int32_t)cur[A]);
                    lists[id].b.push_back((int32_t)cur[B]);
                }
        }
        const int64_t npair = (int64_t)lists.size();
        std::vector<int32_t> pA(npair), pB(npair), er, ea, eb;
        std::vector<int64_t> poff(npair + 1, 0), vpoff(npair + 1, 0);
        for (int64_t q = 0; q < npair; ++q) {
            pA[q] = pair_atoms[q].first;
            pB[q] = pair_atoms[q].second;
            poff[q + 1] = poff[q] + (int64_t)lists[q].r.size();
            vpoff[q + 1] = vpoff[q] + (int64_t)nf[pA[q]] * nf[pB[q]];
        }
        er.reserve(poff[npair]);
        ea.reserve(poff[npair]);
        eb.reserve(poff[npair]);
        for (auto& L : lists) {
            er.insert(er.end(), L.r.begin(), L.r.end());
            ea.insert(ea.end(), L.a.begin(), L.a.end());
            eb.insert(eb.end(), L.b.begin(), L.b.end());
        }
        // per window: dense T_W and, per entry, the place of its V value in vpair
        // (-1: the two functions' atoms share no grid point), both in the T-last order
        require(vpoff[npair] < (int64_t(1) << 31), "pair blocks exceed 2^31 values");
        std::vector<int32_t> lat(Woff[nwin]), lf(Woff[nwin]), pairtab;
        std::vector<double> TW(P->d_SW.n, 0.0);
        std::vector<int32_t> hmap(P->d_SW.n, -1);
        std::vector<int32_t> pos_all(Woff[nwin]);
        hip_check(hipMemcpy(pos_all.data(), P->d_pos.p, pos_all.size() * sizeof(int32_t), hipMemcpyDeviceToHost),
                  "copy");
        std::vector<int32_t> loc_atom(natom, -1), loc(nb, -1);
        for (int64_t w = 0; w < nwin; ++w) {
            const int64_t* W = Wflat + Woff[w];
            const int n = (int)(Woff[w + 1] - Woff[w]);
            require(n == P->n_of[w], "window tables differ from the plan's");
            const int32_t* pos = pos_all.data() + Woff[w];
            std::vector<int32_t> atoms;
            for (int a = 0; a < n; ++a) {
                const int32_t A = (int32_t)aof[W[a]];
                if (loc_atom[A] < 0) {
                    loc_atom[A] = (int32_t)atoms.size();
                    atoms.push_back(A);
                }
                lat[Woff[w] + pos[a]] = loc_atom[A];
                lf[Woff[w] + pos[a]] = fidx[W[a]];
                loc[W[a]] = a;
            }
            const int natw = (int)atoms.size();
            pairtab.clear();
            for (int i = 0; i < natw; ++i)
                for (int j = 0; j < natw; ++j) {
                    const int32_t A = std::min(atoms[i], atoms[j]), B = std::max(atoms[i], atoms[j]);
                    auto it = pid.find((int64_t)A * natom + B);
                    pairtab.push_back(it == pid.end() ? -1 : it->second);
                }
            const int64_t m0 = P->wins[w].mat_off;
            const int32_t* wl = lat.data() + Woff[w];
            const int32_t* wf = lf.data() + Woff[w];
            for (int pa = 0; pa < n; ++pa)
                for (int pb = 0; pb < n; ++pb) {
                    const int32_t q = pairtab[(size_t)wl[pa] * natw + wl[pb]];
                    if (q < 0) continue;
                    // pairs are stored once with A <= B: the other orientation reads the transposed block
                    const int32_t ga = atoms[wl[pa]], gb = atoms[wl[pb]];
                    hmap[m0 + (int64_t)pa * n + pb] =
                        (int32_t)(vpoff[q] + (ga <= gb ? (int64_t)wf[pa] * nf[gb] + wf[pb] : (int64_t)wf[pb] * nf[ga] + wf[pa]));
                }
            for (int a = 0; a < n; ++a)
                for (int64_t p = Tp[W[a]]; p < Tp[W[a] + 1]; ++p) {
                    const int b = loc[Ti[p]];
                    if (b >= 0) TW[m0 + (int64_t)pos[a] * n + pos[b]] = Tv[p];
                }
            for (int32_t A : atoms) loc_atom[A] = -1;
            for (int a = 0; a < n; ++a) loc[W[a]] = -1;
        }
        hipStream_t st = P->st;
        P->d_pairA.upload(pA.data(), npair, st);
        P->d_pairB.upload(pB.data(), npair, st);
        P->d_pent_off.upload(poff.data(), npair + 1, st);
        P->d_ent_r.upload(er.data(), er.size(), st);
        P->d_ent_a.upload(ea.data(), ea.size(), st);
        P->d_ent_b.upload(eb.data(), eb.size(), st);
        P->d_xat_off.upload(xoff.data(), natom + 1, st);
        P->d_nf_of.upload(nf.data(), natom, st);
        P->d_xat.upload(xat.data(), xat.size(), st);
        P->d_vp_off.upload(vpoff.data(), npair + 1, st);
        P->d_vpair.alloc(std::max<int64_t>(1, vpoff[npair]));
        P->d_hmap.upload(hmap.data(), hmap.size(), st);
        P->d_TW.upload(TW.data(), TW.size(), st);
        P->d_vcur.alloc(ng);
        P->d_vref.alloc(ng);
        P->d_edge_part.alloc(2 * kEdgeBlocks);
        // no keep-edge reference until a lambda_min (a NaN bracket proves no decision)
        const double edge0[kEdgeState] = {NAN, NAN, 0.0, 0.0, 0.0, 0.0};
        P->d_edge.upload(edge0, kEdgeState, st);
        P->d_edge_skip.alloc(1);
        hip_check(hipStreamSynchronize(st), "sync");
        P->nf_max = nf_max;
        P->npair = npair;
        P->have_basis = true;
    });
}

int lres_set_potential(lres_plan* P, const double* d_v) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(P->have_basis, "no basis: call lres_plan_set_basis first");
        // no timing events while the stream is being captured into a graph
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(P->st, &cs), "capture status");
        const bool timed = cs == hipStreamCaptureStatusNone;
        if (timed) hip_check(hipEventRecord(P->ev[0], P->st), "event");
        hip_check(launch_pair_v(P->nf_max, (int)P->npair, P->d_pairA.p, P->d_pairB.p, P->d_pent_off.p, P->d_ent_r.p,
                                P->d_ent_a.p, P->d_ent_b.p, P->d_xat_off.p, P->d_nf_of.p, P->d_xat.p, P->d_vp_off.p,
                                d_v, P->dV, P->d_vpair.p, P->st),
                  "pair kernel");
        hip_check(launch_window_hv(P->d_TW.n, P->d_hmap.p, P->d_vpair.p, P->d_TW.p, P->d_HW.p, P->st),
                  "window H kernel");
        if (timed) hip_check(hipEventRecord(P->ev[1], P->st), "event");
        // the potential of this H and its distance from the keep-edge reference
        hip_check(launch_edge_delta(d_v, P->d_vcur.p, P->d_vref.p, P->ng, P->d_edge_part.p, P->d_edge.p, P->st),
                  "keep-edge bound");
        P->have_H = true;
        P->pot_h = true;
    });
}

int lres_plan_set_stream(lres_plan* P, void* stream) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        P->st = (hipStream_t)stream;
    });
}

static_assert(kMixBlocks == LRES_MIX_PARTS, "mixing workspace size");

int lres_diis_mix(const double* d_coef, int nhist, const double* d_Nh, const double* d_Rh, int64_t ng, double mix,
                  double lo, double ntot, double dV, double* d_out, double* d_part, void* stream) {
    return guarded([&] {
        require(nhist >= 1 && ng >= 1, "empty history or grid");
        require(d_coef && d_Nh && d_Rh && d_out && d_part, "null buffer");
        hip_check(launch_diis_mix(d_coef, nhist, d_Nh, d_Rh, ng, mix, lo, ntot, dV, d_out, d_part,
                                  (hipStream_t)stream),
                  "mix kernels");
    });
}

static_assert(kMaxHist == LRES_MAX_HIST, "history rows of the residual workspace");

int lres_lda_potential(const double* d_n, const double* d_vion, const double* d_vh, int64_t ng, double* d_v,
                       void* stream) {
    return guarded([&] {
        require(ng >= 1, "empty grid");
        require(d_n && d_vion && d_vh && d_v, "null buffer");
        hip_check(launch_lda_potential(d_n, d_vion, d_vh, ng, d_v, (hipStream_t)stream), "potential kernel");
    });
}

int lres_scf_residual(const double* d_nout, double scale, const double* d_nin, int64_t ng, int nhist, int slot,
                      double* d_Nh, double* d_Rh, double* d_part, double* d_stats, void* stream) {
    return guarded([&] {
        require(ng >= 1, "empty grid");
        require(nhist >= 1 && nhist <= LRES_MAX_HIST, "history rows must be 1..LRES_MAX_HIST");
        require(slot >= 0 && slot < nhist, "history slot out of range");
        require(d_nout && d_nin && d_Nh && d_Rh && d_part && d_stats, "null buffer");
        hip_check(launch_scf_residual(d_nout, scale, d_nin, ng, nhist, slot, d_Nh, d_Rh, d_part, d_stats,
                                      (hipStream_t)stream),
                  "residual kernels");
    });
}

// Workspace of lres_scf_residual_noalloc: the per-block partial sums of the
// residual's fixed-order reductions, (LRES_MAX_HIST + 1) rows of LRES_MIX_PARTS
// doubles (72 KiB).  Every GPU the library runs on holds its own copy until the
// process ends; calls on one GPU share it.
static __device__ __align__(256) double scratch[(LRES_MAX_HIST + 1) * LRES_MIX_PARTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device (the
// array's name is a host-side stand-in, never a device address)
double* scratch_on_current_device() {
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    require(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        require(p != nullptr, "no device address for the residual workspace");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int lres_scf_residual_noalloc(const double* d_nout, double scale, const double* d_nin, int64_t ng, int nhist,
                              int slot, double* d_Nh, double* d_Rh, double* d_stats, void* stream) {
    return guarded([&] {
        require(ng >= 1, "empty grid");
        require(nhist >= 1 && nhist <= LRES_MAX_HIST,
                "history rows must be 1..LRES_MAX_HIST (the size of the static workspace)");
        require(slot >= 0 && slot < nhist, "history slot out of range");
        require(d_nout && d_nin && d_Nh && d_Rh && d_stats, "null buffer");
        double* part = scratch_on_current_device();
        // k_scf_residual writes every partial that k_rows_total reads (rows 0..nhist,
        // all LRES_MIX_PARTS blocks), so nothing left by an earlier call is read
        hip_check(launch_scf_residual(d_nout, scale, d_nin, ng, nhist, slot, d_Nh, d_Rh, part, d_stats,
                                      (hipStream_t)stream),
                  "residual kernels");
    });
}

int lres_harris_terms(const double* d_n, const double* d_vh, int64_t ng, double* d_part, double* d_out,
                      void* stream) {
    return guarded([&] {
        require(ng >= 1, "empty grid");
        require(d_n && d_vh && d_part && d_out, "null buffer");
        hip_check(launch_harris(d_n, d_vh, ng, d_part, d_out, (hipStream_t)stream), "Harris kernels");
    });
}

int lres_plan_destroy(lres_plan* P) {
    return guarded([&] {
        if (!P) return;
        (void)hipStreamSynchronize(P->st);
        delete P;
    });
}

int lres_plan_sizes(const lres_plan* P, int32_t* n, int32_t* k) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        std::copy(P->n_of.begin(), P->n_of.end(), n);
        std::copy(P->k_of.begin(), P->k_of.end(), k);
    });
}

int lres_plan_buckets(const lres_plan* P, int32_t* out, int cap, int* nbucket) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        *nbucket = (int)P->buckets.size();
        for (int i = 0; i < (int)P->buckets.size() && i < cap; ++i) {
            const auto& b = P->buckets[i];
            int32_t* o = out + 8 * i;
            o[0] = b.nmax;
            o[1] = (int32_t)b.ids.size();
            o[2] = b.shape[0].T;
            o[3] = b.shape[0].G;
            o[4] = b.shape[0].wide;
            o[5] = b.shape[1].T;
            o[6] = b.shape[1].G;
            o[7] = b.shape[1].wide;
        }
    });
}

int lres_plan_set_precision(lres_plan* P, int precision) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(precision == LRES_FP64 || precision == LRES_MIXED, "unknown precision");
        P->precision = precision;
    });
}

int lres_set_hamiltonian_csr(lres_plan* P, const int64_t* Hp, const int64_t* Hi, const double* Hv, int64_t nnz) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(Hp[P->nb] == nnz && Hp[0] == 0, "H indptr inconsistent with nnz");
        P->d_Hp.upload(Hp, P->nb + 1, P->st);
        P->d_Hi.upload(Hi, nnz, P->st);
        P->d_Hv.upload(Hv, nnz, P->st);
        gather_h(P, P->d_Hp.p, P->d_Hi.p, P->d_Hv.p);
    });
}

int lres_set_hamiltonian_csr_device(lres_plan* P, const int64_t* Hp, const int64_t* Hi, const double* Hv) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        gather_h(P, Hp, Hi, Hv);
    });
}

int lres_window_pass(lres_plan* P, double mu, double e_lo, double* d_dens, double* d_NB, double* d_EB,
                     double* h_sums) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        run_pass(P, mu, e_lo, d_dens, d_NB, d_EB, h_sums);
    });
}

int lres_density_last(lres_plan* P, double* d_dens) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(d_dens != nullptr, "null density buffer");
        require(!P->step_queued, "a queued density step is pending: call lres_step_finish first");
        hip_check(launch_density(P->d_wins.p, (int)P->nwin, P->kmax, P->d_D.p, density_rows(P), d_dens, nullptr,
                                 P->st),
                  "density kernel");
    });
}

int lres_inertia(lres_plan* P, const double* sig, int ns, int64_t* h_total, int32_t* d_counts) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        inertia(P, sig, ns, h_total, d_counts);
    });
}

int lres_lambda_min(lres_plan* P, double tol, double* out) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(tol > 0, "tol must be positive");
        require(P->have_H, "no Hamiltonian loaded");
        const int nwin = (int)P->nwin;
        const double width = lmin_width(tol);
        double cell = 0.0;
        if (lmin_on_device(P) && lambda_min_device(P, tol, &cell)) {
            lambda_min_finish(P, cell, cell + width, out);
            return;
        }
        // upper bound: the smallest Rayleigh quotient H_aa / S_aa over all windows
        if ((int64_t)P->d_rq.n < nwin) P->d_rq.alloc(nwin);
        hip_check(launch_diag_ratio(P->d_wins.p, nwin, P->d_SW.p, P->d_HW.p, P->d_rq.p, P->st), "diag kernel");
        std::vector<double> rq(nwin);
        hip_check(hipMemcpyAsync(rq.data(), P->d_rq.p, nwin * sizeof(double), hipMemcpyDeviceToHost, P->st), "copy");
        hip_check(hipStreamSynchronize(P->st), "sync");
        double hi = *std::min_element(rq.begin(), rq.end());
        hi += 1e-12 * (1.0 + std::fabs(hi));
        int64_t cnt = 0;
        inertia(P, &hi, 1, &cnt, nullptr);
        require(cnt > 0, "inertia count at the Rayleigh bound is zero (indefinite S?)");
        double step = 1.0, lo = hi - step;
        for (int it = 0;; ++it) {
            inertia(P, &lo, 1, &cnt, nullptr);
            if (cnt == 0) break;
            require(it < 60, "no lower bound for the window spectra");
            hi = lo;
            step *= 2.0;
            lo -= step;
        }
        // multisection: K interior shifts per launch, one wave per (window, shift).
        // Every round keeps only the windows that count an eigenvalue below its new
        // upper bracket end: only they can hold the minimum, so the set shrinks as the
        // bracket narrows, and K grows as it does (about kWaves waves per launch, at
        // most kMaxK shifts: log2(K + 1) bits of the bracket per round) on the
        // one-wave-per-shift kernels.
        constexpr int kWaves = kLminWaves, kMaxK = kLminMaxK;
        // the blocked inertia kernel of windows above 128 functions factors each
        // (window, shift) in a workgroup of its own: about kBlocks of them per launch
        // (a pass of the whole chip), so bisection while every window takes part and
        // up to kMaxK shifts once a few are left
        constexpr int kBlocks = 512;
        const bool blocked = std::any_of(P->buckets.begin(), P->buckets.end(),
                                         [](const lres_plan::Bucket& b) { return b.big; });
        std::vector<double> sg(kMaxK);
        std::vector<int64_t> tot(kMaxK);
        std::vector<int32_t> cand, next;
        std::vector<int32_t> cnt_w((size_t)nwin * kMaxK);
        if ((int64_t)P->d_icnt.n < (int64_t)nwin * kMaxK) P->d_icnt.alloc((size_t)nwin * kMaxK);
        bool all = true;
        while (hi - lo > tol) {
            const int ncand = all ? nwin : (int)cand.size();
            const int K = blocked ? std::max(1, std::min(kMaxK, kBlocks / std::max(1, ncand)))
                                  : std::max(kLminMinK, std::min(kMaxK, kWaves / std::max(1, ncand)));
            for (int s = 0; s < K; ++s) sg[s] = lo + (hi - lo) * (s + 1) / (K + 1);
            // the per-window counts come back while the set is large enough for
            // shrinking it to pay (small sets run at up to kMaxK shifts anyway), and
            // always for blocked windows (each costs a factorisation per shift)
            const bool shrink = blocked || ncand > 64;
            inertia(P, sg.data(), K, tot.data(), P->d_icnt.p, all ? nullptr : &cand, shrink ? cnt_w.data() : nullptr);
            double nlo = lo, nhi = hi;
            int first_pos = -1;
            for (int s = 0; s < K; ++s) {
                if (tot[s] == 0) nlo = sg[s];
                else { nhi = sg[s]; first_pos = s; break; }
            }
            if (first_pos >= 0 && shrink) {
                next.clear();
                if (all) {
                    for (int w = 0; w < nwin; ++w)
                        if (cnt_w[(size_t)w * K + first_pos] > 0) next.push_back(w);
                } else {
                    for (int32_t w : cand)
                        if (cnt_w[(size_t)w * K + first_pos] > 0) next.push_back(w);
                }
                if (!next.empty()) {
                    cand.swap(next);
                    all = false;
                }
            }
            if (nlo == lo && nhi == hi) break;   // no progress at double resolution
            lo = nlo;
            hi = nhi;
        }
        // the dyadic cell holding lambda_min, whatever path the search took
        const double g = lmin_grid(lo, hi, width, &cell);
        if (!std::isnan(g)) {
            inertia(P, &g, 1, &cnt, nullptr);
            cell = cnt > 0 ? g - width : g;
        }
        lambda_min_finish(P, cell, cell + width, out);
    });
}

int lres_mu_solve(lres_plan* P, double N, double e_lo, double lo, double hi, double f_lo, double f_hi,
                  double xtol, double ftol, double* d_dens, double* h_mu, double* h_info) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(lo < hi, "mu bracket must satisfy lo < hi");
        int evals = 0;
        double sums[2] = {0.0, 0.0};
        double last_mu = NAN, last_f = NAN, last_e = NAN;
        auto f = [&](double mu) {
            run_pass(P, mu, e_lo, d_dens, nullptr, nullptr, sums);
            ++evals;
            last_mu = mu;
            last_f = 2.0 * sums[0] - N;
            last_e = sums[1];
            return last_f;
        };
        double flo = f_lo, fhi = f_hi;
        if (std::isnan(flo)) flo = f(lo);
        if (std::isnan(fhi)) fhi = f(hi);
        lres::BrentRC br;
        double x = 0.0;
        int r = lres::brent_begin(br, lo, hi, flo, fhi, xtol, ftol, 200, &x);
        require(r != lres::kRcFail, "mu bracket does not change sign");
        while (r == lres::kRcEval) r = lres::brent_feed(br, f(x), &x);
        require(r == lres::kRcDone, "mu solve did not converge");
        const double mu = x;
        *h_mu = mu;
        if (h_info) {
            const bool last = mu == last_mu;
            h_info[0] = evals;
            h_info[1] = last ? last_f : (mu == lo ? flo : (mu == hi ? fhi : NAN));
            h_info[2] = last ? 1.0 : 0.0;
            h_info[3] = last ? last_e : NAN;
        }
    });
}

int lres_mu_solve_warm(lres_plan* P, double N, double e_lo, double mu0, double f0, double slope, double ftol,
                       double xtol, double* h_mu, double* h_info) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(h_mu && h_info, "null output");
        double sums[2] = {0.0, 0.0};
        // LRES_TRACE_MU=1: one line per trial pass on stderr (mu, count, tolerance)
        static const bool trace = std::getenv("LRES_TRACE_MU") != nullptr;
        if (trace)
            fprintf(stderr, "[lres mu] warm search from mu %.12f f %.3e slope %.3e ftol %.3e (%s)\n", mu0, f0, slope,
                    ftol, P->precision == LRES_MIXED ? "c64" : "f64");
        lres::WarmSearch s;
        lres::warm_begin(s, N, e_lo, mu0, f0, slope, ftol, xtol);
        while (!lres::warm_terminal(s)) {
            run_pass(P, s.x, e_lo, nullptr, nullptr, nullptr, sums);
            if (trace) fprintf(stderr, "[lres mu]   pass mu %.12f f %.3e\n", s.x, 2.0 * sums[0] - N);
            lres::warm_feed(s, sums[0], sums[1]);
        }
        require(s.phase != lres::kWarmNoConv, "mu solve did not converge");
        const bool found = s.phase == lres::kWarmDone;
        const bool last = found && s.mu == s.last_mu;
        *h_mu = found ? s.mu : NAN;
        h_info[0] = s.evals;
        h_info[1] = last ? s.last_f : NAN;
        h_info[2] = last ? 1.0 : 0.0;
        h_info[3] = last ? s.last_esum : NAN;
        h_info[4] = s.slope;
        h_info[5] = found ? 0.0 : 1.0;
        h_info[6] = last ? s.last_nsum : NAN;
        h_info[7] = s.by_brent ? 1.0 : 0.0;
    });
}

int lres_step_params(lres_plan* P, double N, double mu0, double e_lo, double margin, double keep_tol, double ftol,
                     double xtol, double slope, int lag) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(!P->step_queued, "a queued density step is pending: call lres_step_finish first");
        require(mu0 > e_lo, "mu must lie above e_lo");
        double* q = P->h_pin + lres_plan::kPrmOff;
        q[kStepPrmN] = N;
        q[kStepPrmMu0] = mu0;
        q[kStepPrmElo] = e_lo;
        q[kStepPrmMargin] = margin;
        q[kStepPrmKeepTol] = keep_tol;
        q[kStepPrmFtol] = ftol;
        q[kStepPrmXtol] = xtol;
        q[kStepPrmSlope] = slope;
        q[kStepPrmLag] = lag ? 1.0 : 0.0;
    });
}

int lres_step_launch(lres_plan* P, int depth, double* d_dens) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(P->have_H, "no Hamiltonian loaded");
        require(!P->step_queued, "a queued density step is pending: call lres_step_finish first");
        require(depth >= 1 && depth <= kMaxStepDepth, "depth must be in 1.." + std::to_string(kMaxStepDepth));
        require(d_dens != nullptr, "null density buffer");
        const int fp32 = P->precision == LRES_MIXED ? 1 : 0;
        const bool cap = capturing(P);
        if (cap) {
            require(passes_prepared(P, fp32), "a step is captured before its kernels were tuned: launch one first");
        } else {
            const double* q = P->h_pin + lres_plan::kPrmOff;
            prepare_passes(P, fp32, 0.5 * (q[kStepPrmMu0] + q[kStepPrmElo]), 0.5 * (q[kStepPrmMu0] - q[kStepPrmElo]));
        }
        step_launch(P, depth, d_dens, !cap);
        if (!cap) {
            P->step_queued = true;
            P->step_depth = depth;
            P->step_timed = true;
        }
    });
}

int lres_step_arm(lres_plan* P, int depth) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(!P->step_queued, "a queued density step is pending: call lres_step_finish first");
        require(depth >= 1 && depth <= kMaxStepDepth, "depth must be in 1.." + std::to_string(kMaxStepDepth));
        P->step_queued = true;
        P->step_depth = depth;
        P->step_timed = false;
    });
}

int lres_step_enqueue(lres_plan* P, double N, double mu0, double e_lo, double margin, double keep_tol, double ftol,
                      double xtol, double slope, int lag, int depth, double* d_dens) {
    const int rc = lres_step_params(P, N, mu0, e_lo, margin, keep_tol, ftol, xtol, slope, lag);
    return rc ? rc : lres_step_launch(P, depth, d_dens);
}

int lres_step_finish(lres_plan* P, double* d_dens, double* h_out) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(P->step_queued, "no density step queued");
        require(h_out != nullptr, "null output");
        P->step_queued = false;
        const int fp32 = P->precision == LRES_MIXED ? 1 : 0;
        // later batches serve searches longer than the first batch held: deeper ones
        int depth = P->step_depth;
        const int more = std::max(depth, kMoreStepDepth);
        const lres::StepState* st = reinterpret_cast<const lres::StepState*>(P->h_pin + 2 * lres_plan::kPinSlots);
        int evals_before = 0, final_before = 0, batches = 1;
        bool timed = P->step_timed;
        for (;;) {
            hip_check(hipStreamSynchronize(P->st), "sync");
            // pole-kernel time of the slots that ran: the trials first, then the final pass
            const int ran = timed ? st->s.evals - evals_before : 0;
            for (int i = 0; i < ran && i < depth; ++i) {
                float ms = 0;
                if (hipEventElapsedTime(&ms, P->ev_slot[2 * i], P->ev_slot[2 * i + 1]) == hipSuccess) {
                    P->stat_passes[fp32] += 1;
                    P->stat_ms[fp32] += ms;
                }
            }
            if (timed && st->final_ran > final_before) {
                float ms = 0;
                if (hipEventElapsedTime(&ms, P->ev_slot[2 * depth], P->ev_slot[2 * depth + 1]) == hipSuccess) {
                    P->stat_passes[fp32] += 1;
                    P->stat_ms[fp32] += ms;
                }
            }
            evals_before = st->s.evals;
            final_before = st->final_ran;
            if (lres::warm_terminal(st->s)) break;
            // the search needs more trials than were queued: another batch
            require(++batches <= 64, "device mu search: too many trial passes");
            hip_check(launch_mu_step(P->d_step.p, lres::kStepResume, P->d_circ.p, P->d_pin + 2 * lres_plan::kPinSlots,
                                     P->st),
                      "mu controller");
            depth = more;
            timed = true;
            step_batch(P, depth, d_dens, true);
        }
        const lres::WarmSearch& s = st->s;
        const bool done = s.phase == lres::kWarmDone;
        h_out[0] = s.phase == lres::kWarmDone ? 0.0 : s.phase == lres::kWarmEdge ? 1.0 : 2.0;
        h_out[1] = s.outcome;
        h_out[2] = s.evals + st->final_ran;
        h_out[3] = done ? s.mu : NAN;
        h_out[4] = done ? s.last_f : NAN;
        h_out[5] = done ? s.last_nsum : NAN;
        h_out[6] = done ? s.last_esum : NAN;
        h_out[7] = s.slope;
        h_out[8] = s.by_brent;
        h_out[9] = batches;
        require(s.phase != lres::kWarmNoConv, "mu solve did not converge");
    });
}

int lres_windows(const int64_t* Hp, const int64_t* Hi, const double* Hv, const int64_t* Sp, const int64_t* Si,
                 const double* Sv, const int64_t* Xp, const int64_t* Xi, const double* Xv, const int64_t* Wflat,
                 const int64_t* Woff, int64_t nwin, const int64_t* Oflat, const int64_t* Ooff, int64_t nb, int64_t ng,
                 double dV, double mu, double e_lo, int M, double* h_dens, double* h_NB, double* h_EB) {
    lres_plan* P = nullptr;
    int rc = lres_plan_create(&P, nb, ng, dV, M, LRES_FP64, Wflat, Woff, nwin, Oflat, Ooff, Sp, Si, Sv, Xp, Xi, Xv,
                              nullptr);
    if (rc) return rc;
    rc = guarded([&] {
        const int64_t nnz = Hp[nb];
        require(lres_set_hamiltonian_csr(P, Hp, Hi, Hv, nnz) == 0, g_err);
        DevBuf<double> dens;
        if (h_dens) {
            dens.alloc(ng);
            hip_check(hipMemsetAsync(dens.p, 0, ng * sizeof(double), P->st), "memset");
        }
        double sums[2];
        run_pass(P, mu, e_lo, h_dens ? dens.p : nullptr, nullptr, nullptr, sums);
        if (h_dens) hip_check(hipMemcpy(h_dens, dens.p, ng * sizeof(double), hipMemcpyDeviceToHost), "copy");
        if (h_NB) hip_check(hipMemcpy(h_NB, P->d_NB.p, nwin * sizeof(double), hipMemcpyDeviceToHost), "copy");
        if (h_EB) hip_check(hipMemcpy(h_EB, P->d_EB.p, nwin * sizeof(double), hipMemcpyDeviceToHost), "copy");
    });
    std::string keep = g_err;
    lres_plan_destroy(P);
    g_err = keep;
    return rc;
}

int lres_plan_layout(const lres_plan* P, int64_t* h_mat_off, int64_t* h_fn) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        for (int64_t w = 0; w < P->nwin; ++w) h_mat_off[w] = P->wins[w].mat_off;
        h_mat_off[P->nwin] = P->nwin ? P->wins[P->nwin - 1].mat_off +
                                           (int64_t)P->wins[P->nwin - 1].n * P->wins[P->nwin - 1].n
                                     : 0;
        if (h_fn) std::copy(P->fn_at.begin(), P->fn_at.end(), h_fn);
    });
}

int lres_set_window_blocks(lres_plan* P, const double* d_HW, const double* d_SW) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        const size_t nn = P->d_SW.n;
        if (d_HW) {
            if (P->d_HW.n < nn) P->d_HW.alloc(nn);
            hip_check(hipMemcpyAsync(P->d_HW.p, d_HW, nn * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
            P->have_H = true;
        }
        if (d_SW) hip_check(hipMemcpyAsync(P->d_SW.p, d_SW, nn * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
        // blocks from elsewhere: no longer H(v) on the plan's S_W
        P->pot_h = false;
        edge_clear(P);
    });
}

int lres_get_window_blocks(const lres_plan* P, double* d_HW, double* d_SW) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        const size_t nn = P->d_SW.n;
        if (d_HW) {
            require(P->have_H, "no Hamiltonian loaded");
            hip_check(hipMemcpyAsync(d_HW, P->d_HW.p, nn * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
        }
        if (d_SW) hip_check(hipMemcpyAsync(d_SW, P->d_SW.p, nn * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
        hip_check(hipStreamSynchronize(P->st), "sync");
    });
}

int lres_pass_blocks(const lres_plan* P, double* d_D, double* d_F) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        const size_t nkk = P->d_Q.n;
        hip_check(launch_mirror(P->d_wins.p, (int)P->nwin, P->d_D.p, P->d_F.p, P->st), "mirror kernel");
        if (d_D) hip_check(hipMemcpyAsync(d_D, P->d_D.p, nkk * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
        if (d_F) hip_check(hipMemcpyAsync(d_F, P->d_F.p, nkk * sizeof(double), hipMemcpyDeviceToDevice, P->st), "copy");
        hip_check(hipStreamSynchronize(P->st), "sync");
    });
}

int lres_adjoint(lres_plan* P, double mu, double e_lo, double alpha, double beta, double* d_gH, double* d_gS) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(P->have_H, "no Hamiltonian loaded: call lres_set_hamiltonian_csr first");
        require(mu > e_lo, "mu must lie above e_lo");
        require(d_gH != nullptr && d_gS != nullptr, "null gradient buffer");
        const double c = 0.5 * (mu + e_lo), R = 0.5 * (mu - e_lo);
        const int64_t half = P->M / 2;
        // windows in index order, in chunks whose two workspaces stay within the budget
        const int64_t budget = int64_t(1) << 28;   // complex entries per workspace (4 GiB each)
        int64_t w0 = 0;
        while (w0 < P->nwin) {
            std::vector<int32_t> ids;
            std::vector<int64_t> off;
            int64_t used = 0;
            int nmax = 0;
            int64_t w = w0;
            for (; w < P->nwin; ++w) {
                const int64_t need = half * P->n_of[w] * P->n_of[w];
                if (!ids.empty() && used + need > budget) break;
                ids.push_back((int32_t)w);
                off.push_back(used);
                used += need;
                nmax = std::max(nmax, (int)P->n_of[w]);
            }
            int bmax = 16;
            while (bmax > 4 && big_lds_bytes(nmax, bmax, 16) > 160 * 1024) bmax /= 2;
            require(big_lds_bytes(nmax, bmax, 16) <= 160 * 1024,
                    "window of " + std::to_string(nmax) + " functions exceeds the adjoint sweep's LDS");
            if ((int64_t)P->d_adjG.n < 2 * used) {
                P->d_adjG.alloc((size_t)(2 * used));
                P->d_adjZ.alloc((size_t)(2 * used));
            }
            P->d_adj_ids.upload(ids.data(), ids.size(), P->st);
            P->d_adj_off.upload(off.data(), off.size(), P->st);
            hip_check(launch_adjoint(bmax, nmax, P->d_wins.p, P->d_adj_ids.p, P->d_adj_off.p, (int)ids.size(), P->d_SW.p,
                                     P->d_HW.p, P->d_Q.p, P->M, c, R, alpha, beta, P->d_adjG.p, P->d_adjZ.p, d_gH,
                                     d_gS, P->st),
                      "adjoint kernels");
            hip_check(hipStreamSynchronize(P->st), "sync");   // the id / offset lists are reused by the next chunk
            w0 = w;
        }
    });
}

int lres_force_setup(lres_plan* P, const double* h_pos, int64_t natom, const int64_t* h_aof, int64_t nb, double rcut,
                     const double* h_exps, int ns, int np, const double* h_box, const int64_t* h_shape, double dx) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(natom > 0 && nb == P->nb, "basis size differs from the plan's");
        const int nf = ns + 3 * np;
        require(nf >= 1 && nf <= 8, "1..8 functions per atom");
        require(nb == natom * nf, "every atom must carry the same functions");
        for (int64_t f = 0; f < nb; ++f) require(h_aof[f] == f / nf, "functions must be atom-major (atom_of_fn)");
        const int rad = (int)std::ceil(rcut / dx);
        const int nbox = 2 * rad + 1;
        for (int a = 0; a < 3; ++a) require(h_shape[a] > nbox, "grid shorter than an atom's box along an axis");
        require((int64_t)h_shape[0] * h_shape[1] * h_shape[2] == P->ng, "grid size differs from the plan's");
        P->f_nf = nf;
        P->f_nbox = nbox;
        P->f_natom = natom;
        for (int a = 0; a < 3; ++a) P->f_shape[a] = h_shape[a];
        // box origins: round(R / dx) - rad, as the basis places its functions
        std::vector<int64_t> org(3 * natom);
        for (int64_t A = 0; A < natom; ++A)
            for (int a = 0; a < 3; ++a) org[3 * A + a] = (int64_t)std::llrint(h_pos[3 * A + a] / dx) - rad;
        auto overlap = [&](int64_t A, int64_t B) {
            for (int a = 0; a < 3; ++a) {
                const int64_t n = h_shape[a];
                int64_t d = (org[3 * B + a] - org[3 * A + a]) % n;
                if (d < 0) d += n;
                if (!(d < nbox || n - d < nbox)) return false;
            }
            return true;
        };
        // force pairs (A <= B, boxes meeting, both in some window) and their window entries
        std::unordered_map<int64_t, int32_t> pid;
        std::vector<int32_t> pA, pB;
        std::vector<std::vector<std::pair<int64_t, int32_t>>> ents;
        int64_t wo = 0;
        for (int64_t w = 0; w < P->nwin; ++w) {
            const int n = P->n_of[w];
            std::vector<std::pair<int64_t, int>> at;   // (atom, first position)
            for (int p = 0; p < n; p += nf) {
                const int64_t fn = P->fn_at[wo + p];
                require(fn % nf == 0 && P->fn_at[wo + p + nf - 1] == fn + nf - 1,
                        "window layout does not keep an atom's functions together");
                at.push_back({fn / nf, p});
            }
            for (const auto& x : at)
                for (const auto& y : at) {
                    if (x.first > y.first || !overlap(x.first, y.first)) continue;
                    const int64_t key = x.first * natom + y.first;
                    auto it = pid.find(key);
                    int32_t id;
                    if (it == pid.end()) {
                        id = (int32_t)pA.size();
                        pid.emplace(key, id);
                        pA.push_back((int32_t)x.first);
                        pB.push_back((int32_t)y.first);
                        ents.emplace_back();
                    } else {
                        id = it->second;
                    }
                    ents[id].push_back({P->wins[w].mat_off + (int64_t)x.second * n + y.second, n});
                }
            wo += n;
        }
        const int64_t npair = (int64_t)pA.size();
        std::vector<int64_t> eptr(npair + 1, 0), eoff;
        std::vector<int32_t> en;
        for (int64_t q = 0; q < npair; ++q) {
            for (const auto& e : ents[q]) {
                eoff.push_back(e.first);
                en.push_back(e.second);
            }
            eptr[q + 1] = (int64_t)eoff.size();
        }
        // per atom: its pair parts in pair order (offset 6 pair + 3 side into the pair output)
        std::vector<std::vector<int64_t>> per(natom);
        for (int64_t q = 0; q < npair; ++q) {
            per[pA[q]].push_back(6 * q);
            if (pB[q] != pA[q]) per[pB[q]].push_back(6 * q + 3);
        }
        std::vector<int64_t> aptr(natom + 1, 0), aent;
        for (int64_t A = 0; A < natom; ++A) {
            aent.insert(aent.end(), per[A].begin(), per[A].end());
            aptr[A + 1] = (int64_t)aent.size();
        }
        // grid owners and owned-CSR rows; T-local position of each atom per window
        std::vector<int64_t> own_g(P->ng);
        hip_check(hipMemcpy(own_g.data(), P->d_own_g.p, P->ng * sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
        std::vector<int32_t> owner(P->ng, -1), row(P->ng, -1);
        for (int64_t w = 0; w < P->nwin; ++w)
            for (int32_t r = 0; r < P->wins[w].nown; ++r) {
                const int64_t g = own_g[P->wins[w].own_off + r];
                owner[g] = (int32_t)w;
                row[g] = (int32_t)(P->wins[w].own_off + r);
            }
        for (int64_t g = 0; g < P->ng; ++g) require(owner[g] >= 0, "grid point owned by no window");
        std::vector<int32_t> tpos((size_t)P->nwin * natom, -1);
        wo = 0;
        for (int64_t w = 0; w < P->nwin; ++w) {
            const int n = P->n_of[w], k = P->k_of[w];
            for (int p = n - k; p < n; p += nf) tpos[(size_t)w * natom + P->fn_at[wo + p] / nf] = p - (n - k);
            wo += n;
        }
        auto up64 = [&](DevBuf<int64_t>& b, const std::vector<int64_t>& h) {
            b.upload(h.empty() ? nullptr : h.data(), h.size(), P->st);
            if (h.empty()) b.alloc(1);
        };
        auto up32 = [&](DevBuf<int32_t>& b, const std::vector<int32_t>& h) {
            b.upload(h.empty() ? nullptr : h.data(), h.size(), P->st);
            if (h.empty()) b.alloc(1);
        };
        up64(P->d_forigin, org);
        up64(P->d_fent_ptr, eptr);
        up64(P->d_fent_off, eoff);
        up64(P->d_faptr, aptr);
        up64(P->d_faent, aent);
        up32(P->d_fpA, pA);
        up32(P->d_fpB, pB);
        up32(P->d_fent_n, en);
        up32(P->d_fowner, owner);
        up32(P->d_frow, row);
        up32(P->d_ftpos, tpos);
        P->d_fpos.upload(h_pos, 3 * natom, P->st);
        P->d_fexps.upload(h_exps, ns + np, P->st);
        P->d_fL.upload(h_box, 3, P->st);
        const size_t nb3 = (size_t)nbox * nbox * nbox;
        P->d_fX.alloc((size_t)natom * nf * nb3);
        P->d_fDX.alloc((size_t)natom * 3 * nf * nb3);
        P->d_fTX.alloc((size_t)natom * nf * nb3);
        P->d_fGH.alloc(std::max<size_t>(1, (size_t)npair * nf * nf));
        P->d_fGS.alloc(std::max<size_t>(1, (size_t)npair * nf * nf));
        P->d_fpair.alloc(std::max<size_t>(1, (size_t)npair * 6));
        P->d_fq.alloc((size_t)natom * 3);
        P->f_npair = npair;
        hip_check(launch_atom_boxes((int)natom, P->d_fpos.p, P->d_forigin.p, nbox, dx, h_box, rcut, P->d_fexps.p, ns,
                                   np, P->dV, P->d_fX.p, P->d_fDX.p, P->d_fTX.p, P->st),
                  "basis box kernel");
        hip_check(hipStreamSynchronize(P->st), "sync");   // host tables above are released on return
        P->have_forces = true;
    });
}

int lres_band_forces(lres_plan* P, const double* d_gH, const double* d_gS, const double* d_D, const double* d_F,
                     double beta, const double* d_v, double* d_forces) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(P->have_forces, "no force set-up: call lres_force_setup first");
        require(d_gH && d_gS && d_D && d_F && d_v && d_forces, "null buffer");
        const int nf = P->f_nf;
        hip_check(launch_force_gather((int)P->f_npair, nf, P->d_fent_ptr.p, P->d_fent_off.p, P->d_fent_n.p, d_gH,
                                      d_gS, P->d_fGH.p, P->d_fGS.p, P->st),
                  "force gather kernel");
        hip_check(launch_force_pairs((int)P->f_npair, nf, P->d_fpA.p, P->d_fpB.p, P->d_forigin.p, P->f_nbox,
                                     P->f_shape, P->d_fX.p, P->d_fDX.p, P->d_fTX.p, d_v, P->d_fGH.p, P->d_fGS.p,
                                     P->d_fpair.p, P->st),
                  "force pair kernel");
        hip_check(launch_force_q((int)P->f_natom, nf, P->d_forigin.p, P->f_nbox, P->f_shape, P->d_fX.p, P->d_fDX.p,
                                 P->d_fowner.p, P->d_frow.p, P->d_ftpos.p, P->d_wins.p, P->d_xptr.p, P->d_xidx.p,
                                 P->d_xval.p, d_D, d_F, beta, P->d_fq.p, P->st),
                  "force Q kernel");
        hip_check(launch_force_total((int)P->f_natom, P->d_faptr.p, P->d_faent.p, P->d_fpair.p, P->d_fq.p, P->dV,
                                     d_forces, P->st),
                  "force total kernel");
    });
}

int lres_plan_stats(lres_plan* P, double* out4, int reset) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        out4[0] = (double)P->stat_passes[0];
        out4[1] = P->stat_ms[0];
        out4[2] = (double)P->stat_passes[1];
        out4[3] = P->stat_ms[1];
        if (reset) {
            P->stat_passes[0] = P->stat_passes[1] = 0;
            P->stat_ms[0] = P->stat_ms[1] = 0.0;
        }
    });
}

int lres_plan_edge_stats(lres_plan* P, double* h_out2) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        require(h_out2 != nullptr, "null output");
        h_out2[0] = h_out2[1] = 0.0;
        if (P->d_edge.n < kEdgeState) return;
        double e[kEdgeState];
        hip_check(hipMemcpyAsync(e, P->d_edge.p, sizeof(e), hipMemcpyDeviceToHost, P->st), "copy");
        hip_check(hipStreamSynchronize(P->st), "sync");
        h_out2[0] = e[4];
        h_out2[1] = e[3];
    });
}

int lres_last_timing(const lres_plan* P, double* ms3) {
    return guarded([&] {
        require(P != nullptr, "null plan");
        ms3[0] = P->timing[0];
        ms3[1] = P->timing[1];
        ms3[2] = P->timing[2];
    });
}

}  // extern "C"
