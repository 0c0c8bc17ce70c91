// C ABI of libastrea_hip.so (declared in include/astrea_hip.h): plans,
// workspace layout, the Runge-Kutta stage programs and the launch sequence of
// one time step.  No HIP allocation happens after ah_plan_create.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/astrea_hip.h"
#include "ah_control.h"
#include "ah_ic.h"
#include "ah_launch.h"
#include "ah_output.h"

using namespace ah;

namespace {

// A tile list the kernels walk with a grid-stride loop; the count lives on
// the device so a captured graph stays valid when the list is compacted.
struct TileList {
    std::vector<Tile> full;
    Tile* d = nullptr;
    int* d_count = nullptr;
    int count = 0;
};

struct Stage {
    int in;  // register holding the stage input (0 = state)
    int op;  // RkOp applied after the stage's L
};

// evolvers.py:79-206, one entry per evolve_space call of the reference
std::vector<Stage> rk_program(int rk) {
    switch (rk) {
        case AH_EULER: return {{0, OP_EULER}};
        case AH_SSPRK22: return {{0, OP_22_0}, {1, OP_22_1}};
        case AH_SSPRK33: return {{0, OP_33_0}, {1, OP_33_1}, {2, OP_33_2}};
        case AH_SSPRK43: return {{0, OP_43_0}, {1, OP_43_1}, {2, OP_43_2}, {3, OP_43_3}};
        case AH_RK4: return {{0, OP_RK4_0}, {1, OP_RK4_1}, {2, OP_RK4_2}, {3, OP_RK4_3}};
        case AH_SSPRK53: return {{0, OP_53_0}, {1, OP_53_1}, {2, OP_53_2}, {3, OP_53_3}, {4, OP_53_4}};
        case AH_SSPRK54: return {{0, OP_54_0}, {1, OP_54_1}, {2, OP_54_2}, {3, OP_54_3}, {4, OP_54_4}};
        case AH_SSPRK104: {
            // the reference evaluates evolve_space 11 times per step (evolvers.py:87-103)
            std::vector<Stage> p = {{0, OP_104_FIRST}};
            for (int i = 0; i < 4; ++i) p.push_back({1, OP_104_ACC1});
            p.push_back({1, OP_104_K5});
            p.push_back({3, OP_104_K2FIRST});
            for (int i = 0; i < 3; ++i) p.push_back({2, OP_104_ACC2});
            p.push_back({2, OP_104_FINAL});
            return p;
        }
    }
    return {};
}

// the one-launch 2D sweep: hydrodynamic PLM/PPM with LLF or HLLC
bool fused_supported(const ah_scheme* sch) {
    return sch->dims == 2 && sch->varmask == AH_VARS_H2 && !sch->magnetic_2d && sch->tile_s == 0 &&
           sch->ppm_author == AH_PPM_MC && !sch->av && (sch->recon == AH_PLM || sch->recon == AH_PPM) &&
           (sch->solver == AH_LLF || sch->solver == AH_HLLC);
}

// PPM authors c/ph (ppm.py:43-62): 1D and 2D hydrodynamics (whose zero
// components make the extrapolant limiter's grid-wide test always true), FP64,
// without the flattener and artificial viscosity the reference ties to author
// 'mc', on the default tiles
bool author_supported(const ah_scheme* sch) {
    return (sch->ppm_author == AH_PPM_C || sch->ppm_author == AH_PPM_PH) &&
           ((sch->dims == 1 && sch->varmask == AH_VARS_H1) || (sch->dims == 2 && sch->varmask == AH_VARS_H2)) &&
           sch->recon == AH_PPM && sch->precision == AH_FP64 && sch->flattener == AH_FLAT_OFF && !sch->av &&
           !sch->magnetic_2d && !sch->fuse2d && sch->tile_s == 0;
}

int popcount8(int m) {
    int n = 0;
    for (int c = 0; c < 8; ++c) n += (m >> c) & 1;
    return n;
}

#define HIPCHK(x)                                  \
    do {                                           \
        if ((x) != hipSuccess) return AH_ERR_HIP;  \
    } while (0)

}  // namespace

struct ah_plan {
    ah_scheme sch;
    int nprob = 0;
    std::vector<ah_problem> probs;
    int nv = 0;
    int varidx[8];
    long long total = 0;
    std::vector<long long> offs;
    KernelSet ks;
    std::vector<Stage> prog;
    int nregs = 0;  // K registers + saved L registers used by the program
    // device tables
    ProbDesc* d_pd = nullptr;
    ProbState* d_ps = nullptr;
    unsigned* d_done = nullptr;  // per problem: tiles of the step's last combination done
    bool fold_finish = false;    // the last combination ends the step (few tiles per problem)
    // tile lists: 0 = 1D sweep, 1/2 = recon ax0/ax1, 3/4 = flux ax0/ax1, 5 = cells (combine),
    // 6/7 = transverse reconstruction ax0/ax1 (MHD)
    TileList tl[8];
    int max_blocks = 8192;  // grid cap per launch (workgroups loop over the tiles beyond it)
    // captured step graphs (ah_plan_step_graph)
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    // snapshots taken inside the step program (ah_plan_set_capture); ctl null: off
    CaptureArgs cap = {};
    // bound memory
    double* U = nullptr;
    double* regs[9] = {nullptr};
    double* Lbuf = nullptr;
    double* Dbuf = nullptr;
    void* wL = nullptr;
    void* wR = nullptr;
    void* bn = nullptr;
    // MHD constrained transport buffers
    void* wF[2] = {nullptr, nullptr};
    void* wD[2] = {nullptr, nullptr};
    void* wU[2] = {nullptr, nullptr};
    double* e3 = nullptr;
    double* scratch = nullptr;
    double* bsave = nullptr;  // B_x, B_y of the state at the start of the step (MHD)
    bool bound = false;
    // kernel timing (ah_plan_profile)
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_kind;
    int ev_used = 0;
};

namespace {
// Tile lists in XCD order.  Workgroup b runs on XCD b % 8 (workgroups are
// dispatched round-robin over the eight XCDs) and walks list positions b,
// b + grid, ...; with a grid that is a multiple of 8, the positions equal to
// x modulo 8 are therefore XCD x's.  Giving each XCD a contiguous band of the
// row-major tile order keeps the halo rows that neighbouring tiles share in
// one L2 instead of fetching them into several.
void xcd_order(std::vector<Tile>& v) {
    constexpr size_t NX = 8;
    const size_t n = v.size();
    if (n < 2 * NX) return;
    const size_t band = (n + NX - 1) / NX;
    std::vector<Tile> o;
    o.reserve(n);
    for (size_t j = 0; j < band; ++j)
        for (size_t x = 0; x < NX; ++x) {
            const size_t s = x * band + j;
            if (s < n) o.push_back(v[s]);
        }
    v.swap(o);
}

struct ProfScope {
    ah_plan* p;
    hipStream_t s;
    bool on;
    ProfScope(ah_plan* p_, int kind, hipStream_t s_) : p(p_), s(s_), on(false) {
        if (p->ev.empty() || 2 * (p->ev_used + 1) > (int)p->ev.size()) return;
        on = hipEventRecord(p->ev[2 * p->ev_used], s) == hipSuccess;
        if (on) p->ev_kind[p->ev_used] = kind;
    }
    ~ProfScope() {
        if (on && hipEventRecord(p->ev[2 * p->ev_used + 1], s) == hipSuccess) p->ev_used++;
    }
};
}  // namespace

namespace {
template <int MASK>
void launch_snapshot(const ah_plan* p, int prob, double* out, hipStream_t s, const CaptureArgs& cap) {
    const ah_problem& q = p->probs[prob];
    const int n1 = p->sch.dims == 1 ? 1 : q.n1;
    const int bc = q.boundary == AH_WRAP ? WRAP : EDGE;
    const bool ho = p->sch.recon == AH_PPM || p->sch.recon == AH_WENO3 || p->sch.recon == AH_WENO5 ||
                    p->sch.recon == AH_WENO7;
    const long long cells = (long long)q.n0 * n1;
    const dim3 g((unsigned)((cells + 255) / 256)), b(256);
    const long long off = p->offs[prob];
    if (p->sch.dims == 1) {
        if (ho) hipLaunchKernelGGL((snapshot_kernel<MASK, 1, true>), g, b, 0, s, p->U, p->total, off, q.n0, n1, bc, q.gamma, out, cap);
        else hipLaunchKernelGGL((snapshot_kernel<MASK, 1, false>), g, b, 0, s, p->U, p->total, off, q.n0, n1, bc, q.gamma, out, cap);
    } else {
        if (ho) hipLaunchKernelGGL((snapshot_kernel<MASK, 2, true>), g, b, 0, s, p->U, p->total, off, q.n0, n1, bc, q.gamma, out, cap);
        else hipLaunchKernelGGL((snapshot_kernel<MASK, 2, false>), g, b, 0, s, p->U, p->total, off, q.n0, n1, bc, q.gamma, out, cap);
    }
}

bool snapshot_any(const ah_plan* p, int prob, double* out, hipStream_t s, const CaptureArgs& cap) {
    switch (p->sch.varmask) {
        case AH_VARS_H1: launch_snapshot<AH_VARS_H1>(p, prob, out, s, cap); return true;
        case AH_VARS_H2: launch_snapshot<AH_VARS_H2>(p, prob, out, s, cap); return true;
        case AH_VARS_M6: launch_snapshot<AH_VARS_M6>(p, prob, out, s, cap); return true;
        case AH_VARS_ALL: launch_snapshot<AH_VARS_ALL>(p, prob, out, s, cap); return true;
    }
    return false;
}
// per-problem status written in stream order into device memory (ah_plan_status_async)
__global__ void status_kernel(const ProbState* ps, int nprob, ah_problem_status* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    ah_problem_status o;
    o.t = ps[p].t;
    o.dt = ps[p].dt;
    o.steps = ps[p].steps;
    o.active = ps[p].active;
    o.failed = ps[p].failed;
    out[p] = o;
}
}  // namespace

extern "C" {

int ah_version(int* major, int* minor) {
    if (major) *major = AH_VERSION_MAJOR;
    if (minor) *minor = AH_VERSION_MINOR;
    return AH_OK;
}

const char* ah_status_str(int status) {
    switch (status) {
        case AH_OK: return "ok";
        case AH_ERR_INVALID: return "invalid argument or option combination";
        case AH_ERR_UNSUPPORTED: return "option combination not implemented on the device path";
        case AH_ERR_HIP: return "HIP runtime error";
        case AH_ERR_NOMEM: return "workspace too small or allocation failed";
        case AH_ERR_STATE: return "plan not bound or used out of order";
        case AH_ERR_TOO_LARGE: return "problem exceeds the static scratch array (AH_STATIC_SCRATCH_BYTES)";
    }
    return "unknown status";
}

int ah_plan_create(const ah_scheme* sch, int nprob, const ah_problem* probs, ah_plan** out) {
    if (!sch || !probs || !out || nprob <= 0) return AH_ERR_INVALID;
    if (sch->dims != 1 && sch->dims != 2) return AH_ERR_INVALID;
    // artificial viscosity: the reference computes it for PPM only (ppm.py:94-95),
    // and its 2D form cannot run (ppm.py:149-154 adds arrays of different shapes)
    if (sch->av && (sch->dims != 1 || sch->recon != AH_PPM)) return AH_ERR_UNSUPPORTED;
    if (sch->magnetic_2d && (sch->dims != 2 || !(sch->varmask & 0x60) || sch->flattener != AH_FLAT_OFF))
        return AH_ERR_UNSUPPORTED;
    ah_plan* p = new (std::nothrow) ah_plan();
    if (!p) return AH_ERR_NOMEM;
    p->sch = *sch;
    p->nprob = nprob;
    p->probs.assign(probs, probs + nprob);
    p->nv = popcount8(sch->varmask);
    for (int c = 0, k = 0; c < 8; ++c)
        if ((sch->varmask >> c) & 1) p->varidx[k++] = c;
    bool ok;
    // tile_s: the 1D sweep's tile length (0: the default); 115 selects the
    // alternative shape of 1D hydrodynamic PLM/PPM with LLF/HLLC
    if (sch->tile_s != 0 && (sch->dims != 1 || sch->ppm_author != AH_PPM_MC))
        ok = false;
    else if (sch->fuse2d)
        ok = fused_supported(sch) &&
             (sch->precision == AH_FP64 ? kernels_2d_fused_fp64(sch->recon, sch->solver, &p->ks)
                                        : kernels_2d_fused_fp32(sch->recon, sch->solver, &p->ks));
    else if (sch->tile_s != 0)
        ok = (sch->precision == AH_FP64 ? kernels_1d_small_fp64(sch->varmask, sch->recon, sch->solver, &p->ks)
                                        : kernels_1d_small_fp32(sch->varmask, sch->recon, sch->solver, &p->ks)) &&
             p->ks.ts1 == sch->tile_s;
    else if (sch->ppm_author != AH_PPM_MC)
        ok = author_supported(sch) && (sch->dims == 1 ? kernels_1d_author_fp64(sch->ppm_author, sch->solver, &p->ks)
                                                       : kernels_2d_author_fp64(sch->ppm_author, sch->solver, &p->ks));
    else if (sch->precision == AH_FP64)
        ok = sch->dims == 1 ? kernels_1d_fp64(sch->varmask, sch->recon, sch->solver, &p->ks)
                            : kernels_2d_fp64(sch->varmask, sch->recon, sch->solver, &p->ks);
    else
        ok = sch->dims == 1 ? kernels_1d_fp32(sch->varmask, sch->recon, sch->solver, &p->ks)
                            : kernels_2d_fp32(sch->varmask, sch->recon, sch->solver, &p->ks);
    if (ok && sch->magnetic_2d && sch->precision != AH_FP64) {
        // the split path's constrained transport runs the FP64 kernels
        KernelSet ct;
        ok = kernels_2d_fp64(sch->varmask, sch->recon, sch->solver, &ct);
        for (int ax = 0; ax < 2; ++ax) p->ks.trecon[ax] = ct.trecon[ax];
        p->ks.corner = ct.corner;
        p->ks.magL = ct.magL;
        p->ks.bface = ct.bface;
        p->ks.refine = ct.refine;
        p->ks.copyB = ct.copyB;
        p->ks.tts = ct.tts;
        p->ks.ttt = ct.ttt;
    }
    p->prog = rk_program(sch->rk);
    if (sch->magnetic_2d && !p->ks.corner) ok = false;
    if (!ok || p->prog.empty()) {
        delete p;
        return AH_ERR_UNSUPPORTED;
    }
    // registers: K1..K4 (1-4), saved L (5-8)
    switch (sch->rk) {
        case AH_EULER: p->nregs = 0; break;
        case AH_SSPRK22: p->nregs = 1; break;
        case AH_SSPRK33: p->nregs = 2; break;
        case AH_SSPRK43: p->nregs = 3; break;
        case AH_RK4: p->nregs = 7; break;
        case AH_SSPRK53: p->nregs = 6; break;
        case AH_SSPRK54: p->nregs = 5; break;
        case AH_SSPRK104: p->nregs = sch->magnetic_2d ? 4 : 3; break;  // + the refined increment
    }
    // problem tables and tiles
    std::vector<ProbDesc> pd(nprob);
    long long max_ctiles = 0;
    long long total = 0;
    for (int i = 0; i < nprob; ++i) {
        const ah_problem& q = probs[i];
        const int n1 = sch->dims == 1 ? 1 : q.n1;
        if (q.n0 <= 0 || n1 <= 0 || q.dx <= 0 || q.checkpoints <= 0 || (q.boundary != AH_EDGE && q.boundary != AH_WRAP)) {
            delete p;
            return AH_ERR_INVALID;
        }
        pd[i].n0 = q.n0;
        pd[i].n1 = n1;
        pd[i].bc = q.boundary;
        pd[i].nprob_pad = 0;
        pd[i].off = total;
        pd[i].dx = q.dx;
        pd[i].gamma = q.gamma;
        pd[i].cfl = q.cfl;
        pd[i].t_end = q.t_end;
        pd[i].chkpt = q.t_end / q.checkpoints;
        pd[i].max_steps = q.max_steps;
        p->offs.push_back(total);
        const long long cells = (long long)q.n0 * n1;
        total += cells;
        if (sch->dims == 1) {
            for (int s0 = 0; s0 < q.n0; s0 += p->ks.ts1) p->tl[0].full.push_back({i, s0, 0, 0});
        } else {
            for (int ax = 0; ax < 2; ++ax) {
                const int ns = ax == 0 ? q.n0 : n1, nt = ax == 0 ? n1 : q.n0;
                if (p->ks.recon[ax])  // no reconstruction launch when the flux tiles reconstruct
                    for (int t0 = 0; t0 < nt; t0 += p->ks.rtt[ax])
                        for (int s0 = 0; s0 < ns; s0 += p->ks.rts[ax]) p->tl[1 + ax].full.push_back({i, s0, t0, 0});
                for (int t0 = 0; t0 < nt; t0 += p->ks.ftt[ax])
                    for (int s0 = 0; s0 < ns; s0 += p->ks.fts[ax]) p->tl[3 + ax].full.push_back({i, s0, t0, 0});
                if (sch->magnetic_2d)
                    for (int t0 = 0; t0 < nt; t0 += p->ks.ttt)
                        for (int s0 = 0; s0 < ns; s0 += p->ks.tts) p->tl[6 + ax].full.push_back({i, s0, t0, 0});
            }
        }
        // combination tiles: 256 cells (one per thread) in 1D, so the few large
        // problems left at the end of an ensemble still spread over many CUs;
        // 1024 in 2D, where four cells per thread ran faster on 1024^2 grids
        const int ct = sch->dims == 1 ? 256 : 1024;
        for (long long c0 = 0; c0 < cells; c0 += ct)
            p->tl[5].full.push_back({i, (int)c0, (int)std::min<long long>(ct, cells - c0), 0});
        max_ctiles = std::max(max_ctiles, (cells + ct - 1) / ct);
    }
    p->total = total;
    // each combination tile of the step's last combination counts itself on
    // its problem's counter; beyond a few dozen tiles per problem the atomics
    // on one address cost more than the separate end-of-step launch (1024^2:
    // 1024 tiles, +5 us against the launch's 5 us)
    p->fold_finish = !sch->magnetic_2d && max_ctiles <= 64;
    if (total >= (1ll << 31)) {
        delete p;
        return AH_ERR_INVALID;
    }
    for (int l : {0, 1, 2, 3, 4, 6, 7}) xcd_order(p->tl[l].full);
    bool up = hipMalloc((void**)&p->d_pd, nprob * sizeof(ProbDesc)) == hipSuccess &&
              hipMemcpy(p->d_pd, pd.data(), nprob * sizeof(ProbDesc), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc((void**)&p->d_ps, nprob * sizeof(ProbState)) == hipSuccess &&
              hipMemset(p->d_ps, 0, nprob * sizeof(ProbState)) == hipSuccess &&
              hipMalloc((void**)&p->d_done, nprob * sizeof(unsigned)) == hipSuccess &&
              hipMemset(p->d_done, 0, nprob * sizeof(unsigned)) == hipSuccess;
    for (TileList& L : p->tl) {
        if (!up) break;
        L.count = (int)L.full.size();
        up = hipMalloc((void**)&L.d_count, sizeof(int)) == hipSuccess &&
             hipMemcpy(L.d_count, &L.count, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        if (up && L.count)
            up = hipMalloc((void**)&L.d, L.full.size() * sizeof(Tile)) == hipSuccess &&
                 hipMemcpy(L.d, L.full.data(), L.full.size() * sizeof(Tile), hipMemcpyHostToDevice) == hipSuccess;
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    // 32 workgroups per CU: a 1024^2 grid's sweeps then launch one workgroup
    // per tile and the hardware balances the last round (2048 persistent
    // workgroups left a partly idle final round: 1.023 -> 0.984 ms per
    // shear-layer step); much larger plans still loop over several tiles
    p->max_blocks = 32 * cus;
    if (!up) {
        ah_plan_destroy(p);
        return AH_ERR_HIP;
    }
    *out = p;
    return AH_OK;
}

int ah_plan_destroy(ah_plan* p) {
    if (!p) return AH_ERR_INVALID;
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    int rc = AH_OK;
    if (p->graph_exec && hipGraphExecDestroy(p->graph_exec) != hipSuccess) rc = AH_ERR_HIP;
    void* bufs[] = {p->d_pd, p->d_ps, p->d_done};
    for (void* b : bufs)
        if (b && hipFree(b) != hipSuccess) rc = AH_ERR_HIP;
    for (TileList& L : p->tl) {
        if (L.d && hipFree(L.d) != hipSuccess) rc = AH_ERR_HIP;
        if (L.d_count && hipFree(L.d_count) != hipSuccess) rc = AH_ERR_HIP;
    }
    delete p;
    return rc;
}

int ah_plan_layout(const ah_plan* p, long long* total_cells, int* nvars, int* varidx, long long* offsets) {
    if (!p) return AH_ERR_INVALID;
    if (total_cells) *total_cells = p->total;
    if (nvars) *nvars = p->nv;
    if (varidx)
        for (int v = 0; v < p->nv; ++v) varidx[v] = p->varidx[v];
    if (offsets)
        for (int i = 0; i < p->nprob; ++i) offsets[i] = p->offs[i];
    return AH_OK;
}

static size_t plane_bytes(const ah_plan* p, size_t elem) { return ((size_t)p->nv * p->total * elem + 255) & ~size_t(255); }

int ah_plan_workspace_size(const ah_plan* p, size_t* bytes) {
    if (!p || !bytes) return AH_ERR_INVALID;
    size_t b = plane_bytes(p, 8) * (p->nregs + 1);  // registers + L
    if (p->sch.dims == 2) {
        b += plane_bytes(p, 8);                        // D
        if (p->ks.recon[0]) b += 2 * plane_bytes(p, p->ks.real_bytes);  // wL, wR
        b += ((size_t)p->total * p->ks.real_bytes + 255) & ~size_t(255);  // normal field for HLLD
        if (p->sch.magnetic_2d) {
            b += 6 * plane_bytes(p, p->ks.real_bytes);                          // wF, wD, wU per sweep
            b += 5 * (((size_t)p->total * 8 + 255) & ~size_t(255));            // e3, 2 scratch, 2 B backup planes
        }
    }
    *bytes = b;
    return AH_OK;
}

int ah_plan_bind(ah_plan* p, double* state, void* ws, size_t ws_bytes) {
    if (!p || !state || (!ws && p->nregs + 1 > 0)) return AH_ERR_INVALID;
    size_t need;
    ah_plan_workspace_size(p, &need);
    if (ws_bytes < need) return AH_ERR_NOMEM;
    // a captured step graph holds the previously bound buffers
    if (p->graph_exec) {
        HIPCHK(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
        p->graph_steps = 0;
    }
    char* c = (char*)ws;
    p->U = state;
    p->regs[0] = state;
    for (int r = 1; r <= 8; ++r) p->regs[r] = nullptr;
    // K registers first, then saved L registers, depending on the scheme
    int nk = 0, nl = 0;
    switch (p->sch.rk) {
        case AH_SSPRK22: nk = 1; break;
        case AH_SSPRK33: nk = 2; break;
        case AH_SSPRK43: nk = 3; break;
        case AH_RK4: nk = 3; nl = 3; break;
        case AH_SSPRK53: nk = 4; nl = 2; break;
        case AH_SSPRK54: nk = 4; nl = 1; break;
        case AH_SSPRK104: nk = p->sch.magnetic_2d ? 4 : 3; break;
    }
    for (int r = 0; r < nk; ++r) {
        p->regs[1 + r] = (double*)c;
        c += plane_bytes(p, 8);
    }
    for (int r = 0; r < nl; ++r) {
        p->regs[5 + r] = (double*)c;
        c += plane_bytes(p, 8);
    }
    p->Lbuf = (double*)c;
    c += plane_bytes(p, 8);
    if (p->sch.dims == 2) {
        p->Dbuf = (double*)c;
        c += plane_bytes(p, 8);
        p->wL = p->wR = nullptr;
        if (p->ks.recon[0]) {
            p->wL = c;
            c += plane_bytes(p, p->ks.real_bytes);
            p->wR = c;
            c += plane_bytes(p, p->ks.real_bytes);
        }
        p->bn = c;
        c += ((size_t)p->total * p->ks.real_bytes + 255) & ~size_t(255);
        if (p->sch.magnetic_2d) {
            for (int ax = 0; ax < 2; ++ax) {
                p->wF[ax] = c;
                c += plane_bytes(p, p->ks.real_bytes);
                p->wD[ax] = c;
                c += plane_bytes(p, p->ks.real_bytes);
                p->wU[ax] = c;
                c += plane_bytes(p, p->ks.real_bytes);
            }
            p->e3 = (double*)c;
            c += ((size_t)p->total * 8 + 255) & ~size_t(255);
            p->scratch = (double*)c;
            c += 2 * (((size_t)p->total * 8 + 255) & ~size_t(255));
            p->bsave = (double*)c;
        }
    }
    p->bound = true;
    return AH_OK;
}

int ah_plan_reset(ah_plan* p, void* stream) {
    if (!p) return AH_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(s));
    for (TileList& L : p->tl) {  // every problem is active again: full work lists
        if (L.full.empty()) continue;
        L.count = (int)L.full.size();
        HIPCHK(hipMemcpy(L.d, L.full.data(), L.full.size() * sizeof(Tile), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L.d_count, &L.count, sizeof(int), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(reset_kernel, dim3((p->nprob + 255) / 256), dim3(256), 0, s, p->d_pd, p->d_ps, p->nprob);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

static SweepArgs base_args(ah_plan* p) {
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    a.plane = p->total;
    a.pd = p->d_pd;
    a.ps = p->d_ps;
    a.limiter = p->sch.limiter;
    a.flattener = p->sch.flattener;
    a.parity_override = -1;
    a.compat = p->sch.compat;
    a.av = p->sch.av;
    return a;
}

static int blocks_for(const ah_plan* p, const TileList& L) { return std::min<int>((int)L.full.size(), p->max_blocks); }

// L of the stage input X for every active problem (evolve_space + compute_L)
static void launch_space(ah_plan* p, const double* X, double* Lout, int stage0, int parity_override,
                         double* flux0, double* flux1, hipStream_t s, bool fuse33 = false,
                         const double* xl = nullptr, int stage_idx = 0) {
    SweepArgs a = base_args(p);
    a.X = X;
    a.xl = xl;
    a.stage_idx = stage_idx;
    a.stage0 = stage0;
    a.parity_override = parity_override;
    if (fuse33) {
        a.fuse_k2 = p->regs[2];
        a.fuse_u = p->regs[0];
        a.fuse_k1 = p->regs[1];
    }
    if (p->sch.dims == 1) {
        a.out = Lout;
        a.tiles = p->tl[0].d;
        a.ntiles_dev = p->tl[0].d_count;
        a.flux_out = flux0;
        a.flux_plane = p->probs[0].n0 + 1;
        if (p->sch.ppm_author != AH_PPM_MC) {
            // the interface-limiter tests of this stage over every cell first
            SweepArgs f = a;
            f.flags_only = 1;
            p->ks.sweep1d(f, blocks_for(p, p->tl[0]), s);
        }
        ProfScope ps_(p, AH_K_SWEEP1D, s);
        p->ks.sweep1d(a, blocks_for(p, p->tl[0]), s);
        return;
    }
    a.wL = p->wL;
    a.wR = p->wR;
    a.bn = p->sch.solver == AH_HLLD ? p->bn : nullptr;
    for (int ax = 0; ax < 2; ++ax) {
        a.tiles = p->tl[1 + ax].d;
        a.ntiles_dev = p->tl[1 + ax].d_count;
        a.wF = p->sch.magnetic_2d ? p->wF[ax] : nullptr;
        if (p->ks.recon[ax] && p->sch.ppm_author != AH_PPM_MC) {
            // the interface-limiter tests of this sweep-stage over every cell first
            SweepArgs f = a;
            f.flags_only = 1;
            p->ks.recon[ax](f, blocks_for(p, p->tl[1 + ax]), s);
        }
        if (p->ks.recon[ax]) {
            ProfScope ps_(p, ax == 0 ? AH_K_RECON0 : AH_K_RECON1, s);
            p->ks.recon[ax](a, blocks_for(p, p->tl[1 + ax]), s);
        }
        a.tiles = p->tl[3 + ax].d;
        a.ntiles_dev = p->tl[3 + ax].d_count;
        a.last = ax == 1;
        a.out = ax == 0 ? p->Dbuf : Lout;
        if (ax == 0) a.fuse_k2 = nullptr;  // the first sweep writes its D
        else if (fuse33) a.fuse_k2 = p->regs[2];
        a.Dprev = p->Dbuf;
        a.flux_out = ax == 0 ? flux0 : flux1;
        a.flux_plane = ax == 0 ? (long long)(p->probs[0].n0 + 1) * p->probs[0].n1
                               : (long long)(p->probs[0].n1 + 1) * p->probs[0].n0;
        ProfScope ps_(p, ax == 0 ? AH_K_FLUX0 : AH_K_FLUX1, s);
        p->ks.flux[ax](a, blocks_for(p, p->tl[3 + ax]), s);
    }
}

// stored_dt: combine with the dt stored in the time state (constrained
// transport stores it after stage 0; ah_evolve_time is handed it)
static void launch_combine(ah_plan* p, int op, hipStream_t s, bool stored_dt = false, bool finish = false) {
    CombineArgs c;
    c.finish = finish ? 1 : 0;
    c.ct = p->sch.dims == 1 ? 256 : 1024;
    c.done = p->d_done;
    c.live_dt = (p->sch.magnetic_2d || stored_dt) ? 0 : 1;
    c.dims = p->sch.dims;
    for (int r = 0; r < 9; ++r) c.regs[r] = p->regs[r];
    c.L = p->Lbuf;
    c.pd = p->d_pd;
    c.ps = p->d_ps;
    c.tiles = p->tl[5].d;
    c.ntiles_dev = p->tl[5].d_count;
    c.plane = p->total;
    c.nv = p->nv;
    c.op = op;
    ProfScope ps_(p, AH_K_COMBINE, s);
    hipLaunchKernelGGL(combine_kernel, dim3(blocks_for(p, p->tl[5])), dim3(256), 0, s, c);
}

// register each stage combination writes (for the MHD refinement)
static int target_reg(int op) {
    switch (op) {
        case OP_EULER: case OP_22_1: case OP_33_2: case OP_43_3: case OP_RK4_3: case OP_53_4: case OP_54_4:
        case OP_104_FINAL: return 0;
        case OP_22_0: case OP_33_0: case OP_43_0: case OP_RK4_0: case OP_53_0: case OP_54_0: return 1;
        case OP_33_1: case OP_43_1: case OP_RK4_1: case OP_53_1: case OP_54_1: return 2;
        case OP_43_2: case OP_RK4_2: case OP_53_2: case OP_54_2: case OP_104_K5: return 3;
        case OP_53_3: case OP_54_3: return 4;
    }
    return -1;
}

// SSPRK(10,4) stages that add a refined increment: the addition that
// follows the refinement of K4 = dt/6 L, or -1
static int increment_add(int op) {
    switch (op) {
        case OP_104_FIRST: return OP_104_ADD_U;
        case OP_104_ACC1: return OP_104_ADD_K1;
        case OP_104_K2FIRST: return OP_104_ADD_K3;
        case OP_104_ACC2: return OP_104_ADD_K2;
    }
    return -1;
}

static MhdArgs mhd_args(ah_plan* p);

// A stage's combination followed, with constrained transport, by refine_grid
// (evolvers.py:63-67) of the register it wrote -- or, for the SSPRK(10,4)
// accumulations, of the increment before it is added
static void combine_refine(ah_plan* p, int op, hipStream_t s, bool stored_dt = false, bool finish = false) {
    if (!p->sch.magnetic_2d) {
        launch_combine(p, op, s, stored_dt, finish);
        return;
    }
    MhdArgs m = mhd_args(p);
    const int add = increment_add(op);
    if (add >= 0) {
        launch_combine(p, OP_104_INC, s, stored_dt);
        m.K = p->regs[4];
        p->ks.refine(m, blocks_for(p, p->tl[7]), s);
        p->ks.copyB(m, p->regs[4], blocks_for(p, p->tl[5]), s);
        launch_combine(p, add, s, stored_dt);
        return;
    }
    launch_combine(p, op, s, stored_dt);
    double* K = p->regs[target_reg(op)];
    m.K = K;
    p->ks.refine(m, blocks_for(p, p->tl[7]), s);
    p->ks.copyB(m, K, blocks_for(p, p->tl[5]), s);
}

static MhdArgs mhd_args(ah_plan* p) {
    MhdArgs m;
    memset(&m, 0, sizeof(m));
    for (int ax = 0; ax < 2; ++ax) {
        m.src[ax] = p->wF[ax];
        m.wD[ax] = p->wD[ax];
        m.wU[ax] = p->wU[ax];
        m.ttiles[ax] = p->tl[6 + ax].d;
        m.nttiles_dev[ax] = p->tl[6 + ax].d_count;
    }
    m.e3 = p->e3;
    m.L = p->Lbuf;
    m.U = p->U;
    m.scratch = p->scratch;
    m.bsave = p->bsave;
    m.plane = p->total;
    m.pd = p->d_pd;
    m.ps = p->d_ps;
    m.tiles = p->tl[5].d;
    m.ntiles_dev = p->tl[5].d_count;
    // corner signal speeds (mag_field.py:135-159): the HLL estimate for the
    // HLL family (solver_category 'hll'), the Jacobian's extreme eigenvalues
    // for every other solver
    m.solver_hll = p->sch.solver == AH_HLLC || p->sch.solver == AH_HLLD;
    return m;
}

// one full time step of every active problem
static void launch_step(ah_plan* p, hipStream_t s) {
    const int pb = (p->nprob + 255) / 256;
    const bool mhd = p->sch.magnetic_2d != 0;
    for (size_t st = 0; st < p->prog.size(); ++st) {
        const Stage& S = p->prog[st];
        // the middle SSPRK(3,3) stage combines inside the kernel that finishes L
        // (the 1D sweep, the second 2D flux sweep): dt is known after stage 0,
        // and a failed problem's K2 is never read.  Its input K1 = U + dt L0 is
        // not stored either: the kernels of that stage evaluate it from U and
        // the stage-0 L where they read it (same operations, same bits), so
        // the first stage needs no combination launch
        const bool fuse33 = S.op == OP_33_1 && !mhd;
        // (1D only: in 2D the two reconstruction launches of that stage would
        // each read U and L0 in place of K1, which measured slower than the
        // combination launch it saves -- 0.926 -> 0.946 ms per shear-layer step)
        const bool k1_live = (S.op == OP_33_0 || S.op == OP_33_1) && !mhd && p->sch.dims == 1;
        if (fuse33 && k1_live)
            launch_space(p, p->regs[0], p->Lbuf, 0, -1, nullptr, nullptr, s, true, p->Lbuf, (int)st);
        else
            launch_space(p, p->regs[S.in], p->Lbuf, st == 0, -1, nullptr, nullptr, s, fuse33, nullptr, (int)st);
        MhdArgs m;
        if (mhd) {
            // corner EMF from both sweeps' face values; it replaces the B rows of L
            m = mhd_args(p);
            const int nc = blocks_for(p, p->tl[5]);
            p->ks.trecon[0](m, blocks_for(p, p->tl[6]), s);
            p->ks.trecon[1](m, blocks_for(p, p->tl[7]), s);
            p->ks.corner(m, blocks_for(p, p->tl[7]), s);
            p->ks.magL(m, nc, s);
        }
        // with constrained transport dt is stored after stage 0; otherwise every
        // kernel of the step evaluates it (step_dt) and finish_kernel stores it
        if (st == 0 && mhd) {
            ProfScope ps_(p, AH_K_DT, s);
            hipLaunchKernelGGL(dt_kernel, dim3(pb), dim3(256), 0, s, p->d_pd, p->d_ps, p->nprob, p->sch.dims);
        }
        if (mhd && st == 0) p->ks.bface(m, blocks_for(p, p->tl[7]), s);
        // without constrained transport and with few combination tiles per
        // problem the step's last combination also ends the step
        // (finish_problem in the block that combines a problem's last tile)
        const bool last = st + 1 == p->prog.size() && p->fold_finish;
        if (!fuse33 && !(k1_live && S.op == OP_33_0)) combine_refine(p, S.op, s, false, last);
    }
    if (mhd) {
        ProfScope ps_(p, AH_K_ADVANCE, s);
        hipLaunchKernelGGL(advance_kernel, dim3(pb), dim3(256), 0, s, p->d_pd, p->d_ps, p->nprob);
    } else if (!p->fold_finish) {
        ProfScope ps_(p, AH_K_ADVANCE, s);
        hipLaunchKernelGGL(finish_kernel, dim3(pb), dim3(256), 0, s, p->d_pd, p->d_ps, p->nprob, p->sch.dims);
    }
    if (p->cap.ctl) {
        snapshot_any(p, p->cap.prob, nullptr, s, p->cap);
        hipLaunchKernelGGL(capture_commit_kernel, dim3(1), dim3(64), 0, s, p->cap);
    }
}

}  // extern "C"

namespace {
// The time state of a single-problem plan replaced for a parity entry point
// (active, no failure, the given parity and dt) and put back afterwards.
struct ProbeState {
    ah_plan* p = nullptr;
    hipStream_t s = nullptr;
    ProbState saved;
    int begin(ah_plan* plan, int parity, double dt, hipStream_t stream) {
        p = plan;
        s = stream;
        HIPCHK(hipMemcpyAsync(&saved, p->d_ps, sizeof(ProbState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        ProbState tmp = saved;
        tmp.active = 1;
        tmp.fail_flag = 0;
        tmp.eig[0] = tmp.eig[1] = 0ull;
        tmp.parity = parity;
        tmp.dt = dt;
        tmp.aflags = 0ull;
        HIPCHK(hipMemcpyAsync(p->d_ps, &tmp, sizeof(ProbState), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return AH_OK;
    }
    // `after` receives the state the launches left (eigmax bits) before it is restored
    int end(ProbState* after = nullptr) {
        if (after) {
            HIPCHK(hipMemcpyAsync(after, p->d_ps, sizeof(ProbState), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        HIPCHK(hipMemcpyAsync(p->d_ps, &saved, sizeof(ProbState), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return AH_OK;
    }
};

int slot_of(const ah_plan* p, int comp) {
    for (int v = 0; v < p->nv; ++v)
        if (p->varidx[v] == comp) return v;
    return -1;
}

// the field components the stage-0 reassignment reads (evolvers.py:73-76):
// B_x of the x sweep's face values and B_y of the y sweep's, in the state's
// (x, y) cell order; the sweep-frame arrays are [t][s]
__global__ void ct_faces_kernel(const double* Fx, const double* Fy, double* f0, double* f1, long long plane, int n0,
                                int n1, int s5, int s6) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n0 * n1) return;
    const int i = cell / n1, j = cell % n1;
    f0[cell] = Fx[s5 * plane + (long long)j * n0 + i];
    f1[cell] = Fy[s6 * plane + (long long)i * n1 + j];
}

// B_x, B_y of the state := the given face values; bsave keeps the previous
// ones (the failure path of copyB_kernel restores them)
__global__ void ct_assign_kernel(double* U, double* bsave, const double* f0, const double* f1, long long plane,
                                 int ncell, int s5, int s6) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    bsave[cell] = U[s5 * plane + cell];
    bsave[plane + cell] = U[s6 * plane + cell];
    U[s5 * plane + cell] = f0[cell];
    U[s6 * plane + cell] = f1[cell];
}

void launch_fluxdiff(ah_plan* p, const double* flux0, const double* flux1, double* L, hipStream_t s) {
    const ah_problem& q = p->probs[0];
    const int n1 = p->sch.dims == 1 ? 1 : q.n1;
    const long long ncell = (long long)q.n0 * n1;
    const long long fp0 = (long long)(q.n0 + 1) * n1, fp1 = (long long)(n1 + 1) * q.n0;
    hipLaunchKernelGGL(fluxdiff_kernel, dim3((unsigned)std::min<long long>((ncell + 255) / 256, 4096)), dim3(256), 0,
                       s, flux0, p->sch.dims == 2 ? flux1 : flux0, L, q.n0, n1, p->sch.dims, p->nv, q.dx, p->total,
                       fp0, fp1);
}

// evolve_space of problem 0 (evolvers.py:12-34): fluxes and eigmax per axis;
// with constrained transport also the corner EMF and the face-averaged field
int evolve_space_impl(ah_plan* p, int parity, const double* grid, double* flux0, double* flux1, double* eigmax,
                      double* corner, double* face0, double* face1, hipStream_t s) {
    ProbeState probe;
    int rc = probe.begin(p, parity, 0.0, s);
    if (rc != AH_OK) return rc;
    launch_space(p, grid, p->Lbuf, 1, parity, flux0, flux1, s);
    if (corner) {
        MhdArgs m = mhd_args(p);
        p->ks.trecon[0](m, blocks_for(p, p->tl[6]), s);
        p->ks.trecon[1](m, blocks_for(p, p->tl[7]), s);
        p->ks.corner(m, blocks_for(p, p->tl[7]), s);
        const ah_problem& q = p->probs[0];
        const int ncell = q.n0 * q.n1;
        HIPCHK(hipMemcpyAsync(corner, p->e3, (size_t)ncell * sizeof(double), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(ct_faces_kernel, dim3((ncell + 255) / 256), dim3(256), 0, s, (const double*)p->wF[0],
                           (const double*)p->wF[1], face0, face1, p->total, q.n0, q.n1, slot_of(p, 5), slot_of(p, 6));
    }
    HIPCHK(hipGetLastError());
    ProbState after;
    rc = probe.end(&after);
    if (rc != AH_OK) return rc;
    memcpy(&eigmax[0], &after.eig[0], sizeof(double));
    if (p->sch.dims == 2) memcpy(&eigmax[1], &after.eig[1], sizeof(double));
    return AH_OK;
}
}  // namespace

extern "C" {

int ah_plan_step(ah_plan* p, int nsteps, void* stream) {
    if (!p || nsteps < 0) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    hipStream_t s = (hipStream_t)stream;
    for (int n = 0; n < nsteps; ++n) launch_step(p, s);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_plan_step_graph(ah_plan* p, int nsteps, int nlaunch, void* stream) {
    if (!p || nsteps <= 0 || nlaunch < 0 || !stream) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    hipStream_t s = (hipStream_t)stream;
    if (!p->graph_exec || p->graph_steps != nsteps) {
        if (p->graph_exec) {
            HIPCHK(hipGraphExecDestroy(p->graph_exec));
            p->graph_exec = nullptr;
        }
        std::vector<hipEvent_t> saved;
        saved.swap(p->ev);  // no event records inside a captured graph
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int n = 0; n < nsteps; ++n) launch_step(p, s);
        const hipError_t e1 = hipStreamEndCapture(s, &g);
        saved.swap(p->ev);
        if (e1 != hipSuccess || !g) return AH_ERR_HIP;
        const hipError_t e2 = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            p->graph_exec = nullptr;
            return AH_ERR_HIP;
        }
        p->graph_steps = nsteps;
    }
    for (int i = 0; i < nlaunch; ++i) HIPCHK(hipGraphLaunch(p->graph_exec, s));
    return AH_OK;
}

int ah_plan_init_ic(ah_plan* p, int prob, const ah_ic* ic, void* stream) {
    if (!p || !ic || prob < 0 || prob >= p->nprob) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    if (ic->kind != AH_IC_RIEMANN_1D && ic->kind != AH_IC_QUADRANTS_2D) return AH_ERR_INVALID;
    if ((ic->kind == AH_IC_RIEMANN_1D) != (p->sch.dims == 1)) return AH_ERR_INVALID;
    const ah_problem& q = p->probs[prob];
    IcArgs a;
    memset(&a, 0, sizeof(a));
    a.state = p->U;
    a.plane = p->total;
    a.off = p->offs[prob];
    a.n0 = q.n0;
    a.n1 = p->sch.dims == 1 ? 1 : q.n1;
    a.dims = p->sch.dims;
    a.bc = q.boundary == AH_WRAP ? WRAP : EDGE;
    a.nv = p->nv;
    for (int v = 0; v < p->nv; ++v) a.varidx[v] = p->varidx[v];
    a.kind = ic->kind;
    a.high_order = ic->high_order != 0;
    a.gamma = q.gamma;
    a.x0 = ic->x0;
    a.y0 = ic->y0;
    memcpy(a.w, ic->w, sizeof(a.w));
    const long long cells = (long long)a.n0 * a.n1;
    hipLaunchKernelGGL(ic_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_plan_snapshot(ah_plan* p, int prob, double* out, void* stream) {
    if (!p || !out || prob < 0 || prob >= p->nprob) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    CaptureArgs none;
    memset(&none, 0, sizeof(none));
    if (!snapshot_any(p, prob, out, (hipStream_t)stream, none)) return AH_ERR_UNSUPPORTED;
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_plan_set_capture(ah_plan* p, int prob, int mode, double* ring, int slots, double* times, long long* ctl,
                        void* stream) {
    if (!p || prob < 0 || prob >= p->nprob || mode < AH_CAPTURE_OFF || mode > AH_CAPTURE_CHECKPOINTS)
        return AH_ERR_INVALID;
    if (mode != AH_CAPTURE_OFF && (!ring || !times || !ctl || slots <= 0)) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    // a captured step graph holds the previous capture arguments
    if (p->graph_exec) {
        HIPCHK(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
        p->graph_steps = 0;
    }
    memset(&p->cap, 0, sizeof(p->cap));
    if (mode == AH_CAPTURE_OFF) return AH_OK;
    const ah_problem& q = p->probs[prob];
    p->cap.ps = p->d_ps;
    p->cap.ring = ring;
    p->cap.times = times;
    p->cap.ctl = ctl;
    p->cap.cells = (long long)q.n0 * (p->sch.dims == 1 ? 1 : q.n1);
    p->cap.t_end = q.t_end;
    p->cap.prob = prob;
    p->cap.mode = mode;
    p->cap.slots = slots;
    hipLaunchKernelGGL(capture_init_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, p->cap);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_plan_compact(ah_plan* p, int* nactive, void* stream) {
    if (!p) return AH_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    std::vector<ProbState> h(p->nprob);
    HIPCHK(hipMemcpyAsync(h.data(), p->d_ps, p->nprob * sizeof(ProbState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int na = 0;
    for (const ProbState& q : h) na += q.active ? 1 : 0;
    if (nactive) *nactive = na;
    for (TileList& L : p->tl) {
        if (L.full.empty()) continue;
        std::vector<Tile> act;
        act.reserve(L.full.size());
        for (const Tile& t : L.full)
            if (h[t.prob].active) act.push_back(t);
        L.count = (int)act.size();
        if (L.count) HIPCHK(hipMemcpy(L.d, act.data(), act.size() * sizeof(Tile), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L.d_count, &L.count, sizeof(int), hipMemcpyHostToDevice));
    }
    return AH_OK;
}

int ah_plan_set_blocks(ah_plan* p, int max_blocks) {
    if (!p || max_blocks <= 0) return AH_ERR_INVALID;
    p->max_blocks = max_blocks;
    if (p->graph_exec) {
        HIPCHK(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
    }
    return AH_OK;
}

int ah_plan_profile(ah_plan* p, int capacity) {
    if (!p || capacity < 0) return AH_ERR_INVALID;
    for (hipEvent_t e : p->ev)
        if (hipEventDestroy(e) != hipSuccess) return AH_ERR_HIP;
    p->ev.clear();
    p->ev_kind.clear();
    p->ev_used = 0;
    if (capacity == 0) return AH_OK;
    p->ev.resize(2 * (size_t)capacity);
    p->ev_kind.resize(capacity);
    for (auto& e : p->ev) HIPCHK(hipEventCreate(&e));
    return AH_OK;
}

int ah_plan_profile_read(ah_plan* p, double* ms, int* counts) {
    if (!p || !ms || !counts) return AH_ERR_INVALID;
    for (int k = 0; k < AH_NKIND; ++k) {
        ms[k] = 0.0;
        counts[k] = 0;
    }
    if (p->ev_used > 0) HIPCHK(hipEventSynchronize(p->ev[2 * p->ev_used - 1]));
    for (int i = 0; i < p->ev_used; ++i) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, p->ev[2 * i], p->ev[2 * i + 1]));
        ms[p->ev_kind[i]] += t;
        counts[p->ev_kind[i]] += 1;
    }
    p->ev_used = 0;
    return AH_OK;
}

int ah_plan_time_export(ah_plan* p, ah_time_state* out, void* stream) {
    if (!p || !out) return AH_ERR_INVALID;
    std::vector<ProbState> h(p->nprob);
    HIPCHK(hipMemcpyAsync(h.data(), p->d_ps, p->nprob * sizeof(ProbState), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    for (int i = 0; i < p->nprob; ++i)
        out[i] = ah_time_state{h[i].t, h[i].dt, h[i].steps, h[i].idx, h[i].active, h[i].failed, h[i].parity};
    return AH_OK;
}

int ah_plan_time_import(ah_plan* p, const ah_time_state* in, void* stream) {
    if (!p || !in) return AH_ERR_INVALID;
    std::vector<ProbState> h(p->nprob);
    for (int i = 0; i < p->nprob; ++i) {
        ProbState q;
        memset(&q, 0, sizeof(q));
        q.t = in[i].t;
        q.dt = in[i].dt;
        q.steps = in[i].steps;
        q.idx = in[i].idx;
        q.active = in[i].active;
        q.failed = in[i].failed;
        q.parity = in[i].parity;
        h[i] = q;
    }
    HIPCHK(hipMemcpyAsync(p->d_ps, h.data(), p->nprob * sizeof(ProbState), hipMemcpyHostToDevice, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return AH_OK;
}

int ah_plan_status_async(ah_plan* p, ah_problem_status* out, void* stream) {
    if (!p || !out) return AH_ERR_INVALID;
    hipLaunchKernelGGL(status_kernel, dim3((p->nprob + 255) / 256), dim3(256), 0, (hipStream_t)stream, p->d_ps,
                       p->nprob, out);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_plan_status(ah_plan* p, ah_problem_status* out, void* stream) {
    if (!p || !out) return AH_ERR_INVALID;
    std::vector<ProbState> h(p->nprob);
    HIPCHK(hipMemcpyAsync(h.data(), p->d_ps, p->nprob * sizeof(ProbState), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    for (int i = 0; i < p->nprob; ++i) {
        out[i].t = h[i].t;
        out[i].dt = h[i].dt;
        out[i].steps = h[i].steps;
        out[i].active = h[i].active;
        out[i].failed = h[i].failed;
    }
    return AH_OK;
}

int ah_evolve_space(ah_plan* p, int prob, int parity, const double* grid, double* flux0, double* flux1, double* eigmax,
                    void* stream) {
    if (!p || prob != 0 || p->nprob != 1 || !grid || !flux0 || (p->sch.dims == 2 && !flux1) || !eigmax)
        return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    return evolve_space_impl(p, parity, grid, flux0, flux1, eigmax, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int ah_evolve_space_ct(ah_plan* p, int prob, int parity, const double* grid, double* flux0, double* flux1,
                       double* eigmax, double* corner, double* face0, double* face1, void* stream) {
    if (!p || prob != 0 || p->nprob != 1 || !grid || !flux0 || !flux1 || !eigmax || !corner || !face0 || !face1)
        return AH_ERR_INVALID;
    if (!p->sch.magnetic_2d) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    return evolve_space_impl(p, parity, grid, flux0, flux1, eigmax, corner, face0, face1, (hipStream_t)stream);
}

int ah_compute_L(ah_plan* p, int prob, int parity, const double* flux0, const double* flux1, double* L, void* stream) {
    (void)parity;
    if (!p || prob != 0 || p->nprob != 1 || !flux0 || !L || (p->sch.dims == 2 && !flux1)) return AH_ERR_INVALID;
    // constrained transport overwrites the field rows with EMF differences
    // (evolvers.py:51-58), which need the corner EMF: ah_compute_L_ct
    if (p->sch.magnetic_2d) return AH_ERR_UNSUPPORTED;
    launch_fluxdiff(p, flux0, flux1, L, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return AH_OK;
}

int ah_compute_L_ct(ah_plan* p, int prob, int parity, const double* flux0, const double* flux1, const double* corner,
                    double* L, void* stream) {
    if (!p || prob != 0 || p->nprob != 1 || !flux0 || !flux1 || !corner || !L) return AH_ERR_INVALID;
    if (!p->sch.magnetic_2d) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    hipStream_t s = (hipStream_t)stream;
    ProbeState probe;
    int rc = probe.begin(p, parity, 0.0, s);
    if (rc != AH_OK) return rc;
    launch_fluxdiff(p, flux0, flux1, L, s);
    MhdArgs m = mhd_args(p);
    m.e3 = const_cast<double*>(corner);
    m.L = L;
    p->ks.magL(m, blocks_for(p, p->tl[5]), s);
    HIPCHK(hipGetLastError());
    return probe.end();
}

int ah_evolve_time(ah_plan* p, int prob, int parity, double* grid, const double* flux0, const double* flux1,
                   const double* corner, const double* face0, const double* face1, double dt, void* stream) {
    if (!p || prob != 0 || p->nprob != 1 || !grid || !flux0 || (p->sch.dims == 2 && !flux1)) return AH_ERR_INVALID;
    const bool mhd = p->sch.magnetic_2d != 0;
    if (mhd && (!corner || !face0 || !face1)) return AH_ERR_INVALID;
    if (!p->bound) return AH_ERR_STATE;
    hipStream_t s = (hipStream_t)stream;
    ProbeState probe;
    int rc = probe.begin(p, parity, dt, s);
    if (rc != AH_OK) return rc;
    // the stages update `grid` (the stage registers stay the plan's own)
    double* U_bound = p->U;
    p->U = grid;
    p->regs[0] = grid;
    const ah_problem& q = p->probs[0];
    const int ncell = q.n0 * q.n1;
    // L0 from the given fluxes (evolvers.py:62)
    launch_fluxdiff(p, flux0, flux1, p->Lbuf, s);
    if (mhd) {
        MhdArgs m = mhd_args(p);
        m.e3 = const_cast<double*>(corner);
        p->ks.magL(m, blocks_for(p, p->tl[5]), s);
        // B_x, B_y of the grid := the sweeps' face values, in place (evolvers.py:73-76)
        hipLaunchKernelGGL(ct_assign_kernel, dim3((ncell + 255) / 256), dim3(256), 0, s, grid, p->bsave, face0, face1,
                           p->total, ncell, slot_of(p, 5), slot_of(p, 6));
    }
    for (size_t st = 0; st < p->prog.size(); ++st) {
        const Stage& S = p->prog[st];
        if (st > 0) {
            launch_space(p, p->regs[S.in], p->Lbuf, 0, -1, nullptr, nullptr, s, false, nullptr, (int)st);
            if (mhd) {
                MhdArgs m = mhd_args(p);
                p->ks.trecon[0](m, blocks_for(p, p->tl[6]), s);
                p->ks.trecon[1](m, blocks_for(p, p->tl[7]), s);
                p->ks.corner(m, blocks_for(p, p->tl[7]), s);
                p->ks.magL(m, blocks_for(p, p->tl[5]), s);
            }
        }
        // refine_grid after every stage (evolvers.py:63-67)
        combine_refine(p, S.op, s, true);
    }
    p->U = U_bound;
    p->regs[0] = U_bound;
    HIPCHK(hipGetLastError());
    return probe.end();
}


// ============================================================================
// ah_evolve_space_static: ah_evolve_space for one problem without a plan and
// without allocating.  The problem tables, the tile lists and the interface-
// state workspace live in `scratch`, a fixed array of every device the
// library runs on (held until the process ends).  Calls on one device share
// that array, so calls that overlap on one device race; each call rebuilds
// what it reads (tables, tiles, zeroed eigmax) before its sweeps.
// ============================================================================
}  // extern "C"

static __device__ __align__(256) unsigned char scratch[AH_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};

// the device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side name is not a device address)
int scratch_ptr(unsigned char** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return AH_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return AH_ERR_HIP;
        g_scratch_ptr[dev] = ptr;
    }
    *out = (unsigned char*)g_scratch_ptr[dev];
    return AH_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// tile list k -> (s0 = (k % nsb) * ts, t0 = (k / nsb) * tt) of one problem
struct TileSpec {
    Tile* dst;
    int* count;
    int ns, ts, tt, n;  // n tiles, nsb = ceil(ns / ts)
};

__global__ void static_init_kernel(ProbDesc* pd, ProbDesc desc, ProbState* ps, int parity, TileSpec t0, TileSpec t1,
                                   TileSpec t2, TileSpec t3, int nlists) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        *pd = desc;
        ProbState s;
        s.t = 0.0;
        s.dt = 0.0;
        s.steps = 0;
        s.idx = 1;
        s.active = 1;
        s.failed = 0;
        s.parity = parity;
        s.eig[0] = s.eig[1] = 0ull;
        s.fail_flag = 0;
        s.ckpt = 0;
        s.aflags = 0ull;
        *ps = s;
    }
    const TileSpec specs[4] = {t0, t1, t2, t3};
    for (int l = 0; l < nlists; ++l) {
        const TileSpec& t = specs[l];
        if (k == 0) *t.count = t.n;
        if (k < t.n) {
            const int nsb = (t.ns + t.ts - 1) / t.ts;
            t.dst[k] = Tile{0, (k % nsb) * t.ts, (k / nsb) * t.tt, 0};
        }
    }
}
}  // namespace

extern "C" {

int ah_evolve_space_static(const ah_scheme* sch, const ah_problem* q, int parity, const double* grid, double* flux0,
                           double* flux1, double* eigmax, void* stream) {
    if (!sch || !q || !grid || !flux0 || !eigmax || (sch->dims != 1 && sch->dims != 2) ||
        (sch->dims == 2 && !flux1) || (parity != 0 && parity != 1))
        return AH_ERR_INVALID;
    if (sch->magnetic_2d) return AH_ERR_UNSUPPORTED;
    if (sch->av && (sch->dims != 1 || sch->recon != AH_PPM)) return AH_ERR_UNSUPPORTED;
    if (sch->ppm_author != AH_PPM_MC && !author_supported(sch)) return AH_ERR_UNSUPPORTED;
    const int n0 = q->n0, n1 = sch->dims == 1 ? 1 : q->n1;
    if (n0 <= 0 || n1 <= 0 || q->dx <= 0 || q->checkpoints <= 0 || (q->boundary != AH_EDGE && q->boundary != AH_WRAP))
        return AH_ERR_INVALID;
    KernelSet ks;
    if (sch->tile_s != 0 && (sch->dims != 1 || sch->ppm_author != AH_PPM_MC)) return AH_ERR_UNSUPPORTED;
    if (sch->fuse2d && !fused_supported(sch)) return AH_ERR_UNSUPPORTED;
    bool ok = sch->fuse2d
                  ? (sch->precision == AH_FP64 ? kernels_2d_fused_fp64(sch->recon, sch->solver, &ks)
                                               : kernels_2d_fused_fp32(sch->recon, sch->solver, &ks))
                  : sch->tile_s != 0
                  ? (sch->precision == AH_FP64 ? kernels_1d_small_fp64(sch->varmask, sch->recon, sch->solver, &ks)
                                               : kernels_1d_small_fp32(sch->varmask, sch->recon, sch->solver, &ks)) &&
                        ks.ts1 == sch->tile_s
                  : sch->ppm_author != AH_PPM_MC
                  ? (sch->dims == 1 ? kernels_1d_author_fp64(sch->ppm_author, sch->solver, &ks)
                                    : kernels_2d_author_fp64(sch->ppm_author, sch->solver, &ks))
                  : sch->precision == AH_FP64
                  ? (sch->dims == 1 ? kernels_1d_fp64(sch->varmask, sch->recon, sch->solver, &ks)
                                    : kernels_2d_fp64(sch->varmask, sch->recon, sch->solver, &ks))
                  : (sch->dims == 1 ? kernels_1d_fp32(sch->varmask, sch->recon, sch->solver, &ks)
                                    : kernels_2d_fp32(sch->varmask, sch->recon, sch->solver, &ks));
    if (!ok) return AH_ERR_UNSUPPORTED;
    int nv = 0;
    for (int c = 0; c < 8; ++c) nv += (sch->varmask >> c) & 1;
    const long long cells = (long long)n0 * n1;
    // tile lists: 1D sweep, or recon x/y and flux x/y
    TileSpec spec[4];
    int nlists = 0;
    auto add = [&](int ns, int nt, int ts, int tt) {
        // ts = 0: a list the scheme does not launch (the one-launch 2D sweep
        // has no reconstruction tiles)
        const int n = ts ? ((ns + ts - 1) / ts) * ((nt + tt - 1) / tt) : 0;
        spec[nlists++] = TileSpec{nullptr, nullptr, ns, ts, tt, n};
    };
    if (sch->dims == 1) {
        add(n0, 1, ks.ts1, 1);
    } else {
        add(n0, n1, ks.rts[0], ks.rtt[0]);
        add(n1, n0, ks.rts[1], ks.rtt[1]);
        add(n0, n1, ks.fts[0], ks.ftt[0]);
        add(n1, n0, ks.fts[1], ks.ftt[1]);
    }
    // scratch layout: tables, tile counts and lists, L, D, wL, wR, normal field
    const size_t plane8 = align256((size_t)nv * cells * 8), planeR = align256((size_t)nv * cells * ks.real_bytes);
    size_t need = align256(sizeof(ProbDesc)) + align256(sizeof(ProbState)) + align256(4 * sizeof(int));
    for (int l = 0; l < nlists; ++l) need += align256((size_t)spec[l].n * sizeof(Tile));
    need += plane8;  // L
    if (sch->dims == 2) need += plane8 + (ks.recon[0] ? 2 * planeR : 0) + align256((size_t)cells * ks.real_bytes);
    if (need > (size_t)AH_STATIC_SCRATCH_BYTES || cells >= (1ll << 31)) return AH_ERR_TOO_LARGE;
    unsigned char* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc != AH_OK) return rc;
    unsigned char* c = base;
    ProbDesc* d_pd = (ProbDesc*)c;
    c += align256(sizeof(ProbDesc));
    ProbState* d_ps = (ProbState*)c;
    c += align256(sizeof(ProbState));
    int* d_counts = (int*)c;
    c += align256(4 * sizeof(int));
    for (int l = 0; l < nlists; ++l) {
        spec[l].count = d_counts + l;
        spec[l].dst = (Tile*)c;
        c += align256((size_t)spec[l].n * sizeof(Tile));
    }
    double* Lbuf = (double*)c;
    c += plane8;
    double* Dbuf = nullptr;
    void *wL = nullptr, *wR = nullptr, *bn = nullptr;
    if (sch->dims == 2) {
        Dbuf = (double*)c;
        c += plane8;
        if (ks.recon[0]) {
            wL = c;
            c += planeR;
            wR = c;
            c += planeR;
        }
        bn = c;
    }
    ProbDesc desc;
    memset(&desc, 0, sizeof(desc));
    desc.n0 = n0;
    desc.n1 = n1;
    desc.bc = q->boundary;
    desc.off = 0;
    desc.dx = q->dx;
    desc.gamma = q->gamma;
    desc.cfl = q->cfl;
    desc.t_end = q->t_end;
    desc.chkpt = q->t_end / q->checkpoints;
    desc.max_steps = q->max_steps;
    hipStream_t s = (hipStream_t)stream;
    int maxn = 1;
    for (int l = 0; l < nlists; ++l) maxn = std::max(maxn, spec[l].n);
    for (int l = nlists; l < 4; ++l) spec[l] = spec[0];
    hipLaunchKernelGGL(static_init_kernel, dim3((maxn + 255) / 256), dim3(256), 0, s, d_pd, desc, d_ps, parity, spec[0],
                       spec[1], spec[2], spec[3], nlists);
    HIPCHK(hipGetLastError());
    // the sweeps of launch_space with this problem's tables
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    a.plane = cells;
    a.pd = d_pd;
    a.ps = d_ps;
    a.limiter = sch->limiter;
    a.flattener = sch->flattener;
    a.parity_override = parity;
    a.compat = sch->compat;
    a.av = sch->av;
    a.X = grid;
    a.stage0 = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    auto blocks = [&](int l) { return std::min(spec[l].n, 32 * cus); };
    if (sch->dims == 1) {
        a.out = Lbuf;
        a.tiles = spec[0].dst;
        a.ntiles_dev = spec[0].count;
        a.flux_out = flux0;
        a.flux_plane = n0 + 1;
        if (sch->ppm_author != AH_PPM_MC) {
            SweepArgs f = a;
            f.flags_only = 1;
            ks.sweep1d(f, blocks(0), s);
        }
        ks.sweep1d(a, blocks(0), s);
    } else {
        a.wL = wL;
        a.wR = wR;
        a.bn = sch->solver == AH_HLLD ? bn : nullptr;
        for (int ax = 0; ax < 2; ++ax) {
            a.tiles = spec[ax].dst;
            a.ntiles_dev = spec[ax].count;
            if (ks.recon[ax] && sch->ppm_author != AH_PPM_MC) {
                SweepArgs f = a;
                f.flags_only = 1;
                ks.recon[ax](f, blocks(ax), s);
            }
            if (ks.recon[ax]) ks.recon[ax](a, blocks(ax), s);
            a.tiles = spec[2 + ax].dst;
            a.ntiles_dev = spec[2 + ax].count;
            a.last = ax == 1;
            a.out = ax == 0 ? Dbuf : Lbuf;
            a.Dprev = Dbuf;
            a.flux_out = ax == 0 ? flux0 : flux1;
            a.flux_plane = ax == 0 ? (long long)(n0 + 1) * n1 : (long long)(n1 + 1) * n0;
            ks.flux[ax](a, blocks(2 + ax), s);
        }
    }
    HIPCHK(hipGetLastError());
    ProbState after;
    HIPCHK(hipMemcpyAsync(&after, d_ps, sizeof(ProbState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(&eigmax[0], &after.eig[0], sizeof(double));
    if (sch->dims == 2) memcpy(&eigmax[1], &after.eig[1], sizeof(double));
    return AH_OK;
}
}  // extern "C"
