// batch.hip -- many independent drop simulations advanced together.
//
// Each simulation is a full vdi_ctx (its own fields, gas layer and four
// multigrid hierarchies, of its own grid size and parameters).  A batched
// step replays ONE captured graph whose every node is a persistent kernel
// over all simulations: a work list of (record, tile, simulation) items that
// the grid walks, so the Godunov passes of 64 simulations are one launch,
// level l of 64 MAC V-cycles is one launch, and so on.  Levels a smaller
// grid does not have simply contribute no work items; its coarse solve runs
// in the same launch as the others' (one workgroup per simulation), so the
// down -> coarse -> up order holds for every simulation.
//
// Solves get the per-system budget max over simulations of the cycles the
// previous step needed; converged simulations skip further cycles through
// their device flags.  A simulation that did not converge within the budget
// is restored from the snapshot the graph took and redone alone on its direct
// path, exactly as a single context does.
//
// Tile shapes of the level launches are chosen by vdi_batch_autotune, which
// times every candidate on the batch's own work lists.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>

#include <type_traits>

#include "step.hpp"

namespace vdi {

namespace {

// A batched launch: records and work list in the plan's device blob.
struct BLaunch {
    size_t rec_off = 0, work_off = 0;
    int nw = 0, grid = 8;
    size_t lds = 0;
};

// Host-side builder of the device blob.
struct Blob {
    std::vector<char> bytes;
    size_t put(const void *p, size_t n) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + n);
        if (n) memcpy(bytes.data() + off, p, n);
        return off;
    }
};

// Solve launches of one system kind (k = 0 MAC, 1 VX, 2 VY, 3 FEM).
struct SolvePlan {
    bool used = false, fem = false, mixed = false;
    int maxlc = 0;
    BLaunch extrap, dirichlet, begin, solve_end;
    BLaunch nbegin, ncycle;            // one-workgroup norm reductions (records only)
    std::vector<BLaunch> down, up;     // FP64 path, per level
    BLaunch coarse;
    BLaunch xs[3];                     // multi-workgroup exact solves (members with a larger coarsest level)
    BLaunch down0[2], mx[2];           // mixed path, level 0, by parity
    std::vector<BLaunch> downf, upf;   // mixed path, levels 1..
    BLaunch coarsef, xsf[3];
};

}  // namespace
}  // namespace vdi

using namespace vdi;

struct vdi_batch {
    std::vector<vdi_ctx *> sims;
    int S = 0, form = VDI_FORM_FP64, implicit = 1;
    hipStream_t st = nullptr;
    // tile shape per [kind][prec][which][level]; which 0 down, 1 up, 2 mixed level-0 up
    std::vector<int> tiles[4][2][3];
    DevBuf blob, coef_dev, busy_dev;
    // pinned, mapped, coherent host memory the captured step's kernels read
    // or write directly (the graph holds kernel nodes only): extrapolation
    // coefficients (8 per simulation), readbacks (one per simulation) and the
    // per-simulation clocks (3 per simulation, see busy_end)
    double *coef_host = nullptr, *coef_map = nullptr;
    vdi_ctx::Readback *rb = nullptr, *rb_map = nullptr;
    unsigned long long *busy_host = nullptr, *busy_map = nullptr;
    std::vector<StepRec> recs_host;
    BLaunch step_l[SK_COUNT];
    size_t recs_off = 0;
    SolvePlan solve[4];
    struct GraphEntry {
        int budget[4];
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    int budget[4] = {0, 0, 0, 0};
    long batch_steps = 0, redo_sims = 0, restarts = 0;
    // status of each member's last step (VDI_OK, or why its run ended)
    std::vector<int> status;
    // once every member had its first steps, later steps are always batched
    // (a restarted member's first solves take (1, 0) coefficients, which is
    // what mgs_common::initial_guess does while its history fills)
    bool primed = false;
    double last_ms = 0;
    double ticks_per_ms = 1e5;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool plan_ok = false;
    const char *base = nullptr;  // device blob the launches read (the plan's, or a tuning candidate's)

    void drop_graphs() {
        for (auto &e : graphs) (void)hipGraphExecDestroy(e.exec);
        graphs.clear();
    }
    ~vdi_batch() {
        if (st) (void)hipStreamSynchronize(st);
        drop_graphs();
        for (vdi_ctx *c : sims) vdi_destroy(c);
        if (coef_host) (void)hipHostFree(coef_host);
        if (rb) (void)hipHostFree(rb);
        if (busy_host) (void)hipHostFree(busy_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace vdi {
namespace {

template <class R>
static BLaunch add_launch(Blob &B, const std::vector<R> &recs, const std::vector<WorkItem> &work, size_t lds,
                          const void *kernel) {
    BLaunch L;
    L.rec_off = B.put(recs.data(), recs.size() * sizeof(R));
    L.work_off = B.put(work.data(), work.size() * sizeof(WorkItem));
    L.nw = (int)work.size();
    L.lds = lds;
    L.grid = kernel ? mgk::persistent_grid(kernel, lds, std::max(1, L.nw)) : std::max(8, (L.nw + 7) / 8 * 8);
    return L;
}

static void apply_tiles(vdi_batch *b, int k) {
    for (vdi_ctx *c : b->sims) {
        MG &g = *step_solver(c, k);
        for (int pr = 0; pr < 2; pr++)
            for (int w = 0; w < 2; w++)
                for (int l = 0; l < g.lc; l++) g.tile[pr][w][l] = b->tiles[k][pr][w][l];
        g.tile_mx = b->tiles[k][1][2][0];
    }
}

// Level launch of kind k: which 0 down, 1 up, 2 mixed level-0 up; prec 0/1;
// parity for the mixed level-0 launches.
static BLaunch level_launch(vdi_batch *b, Blob &B, int k, int prec, int which, int l, int parity, double tol2) {
    std::vector<WorkItem> work;
    size_t lds = 0;
    const bool fem = k == 3;
    if (which == 2) {
        std::vector<MxArgs> recs;
        for (int s = 0; s < b->S; s++) {
            MG &g = *step_solver(b->sims[s], k);
            MxArgs a = g.mx_args(parity, tol2);
            for (int t = 0; t < a.nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
            recs.push_back(a);
            lds = std::max(lds, g.mx_lds());
        }
        return add_launch(B, recs, work, lds, mgk::mx_kernel(fem, true));
    }
    if (prec == 0) {
        std::vector<LvArgs<double>> recs;
        for (int s = 0; s < b->S; s++) {
            MG &g = *step_solver(b->sims[s], k);
            if (l >= g.lc) continue;
            LvArgs<double> a = which == 0 ? g.down_args<double>(g.L, l, l > 0) : g.up_args<double>(g.L, l, l == 0, tol2);
            for (int t = 0; t < a.nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
            recs.push_back(a);
            lds = std::max(lds, which == 0 ? g.down_lds(0, l) : g.up_lds(0, l));
        }
        return add_launch(B, recs, work, lds, which == 0 ? mgk::down_kernel(0, fem, true) : mgk::up_kernel(0, fem, true));
    }
    std::vector<LvArgs<float>> recs;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        if (l >= g.lc) continue;
        LvArgs<float> a = which == 0 ? (l == 0 ? g.mx_down0_args(parity) : g.down_args<float>(g.LF, l, 1))
                                     : g.up_args<float>(g.LF, l, 0, tol2);
        for (int t = 0; t < a.nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
        recs.push_back(a);
        lds = std::max(lds, which == 0 ? g.down_lds(1, l) : g.up_lds(1, l));
    }
    return add_launch(B, recs, work, lds, which == 0 ? mgk::down_kernel(1, fem, true) : mgk::up_kernel(1, fem, true));
}

static BLaunch coarse_launch(vdi_batch *b, Blob &B, int k, int prec, double tol2) {
    std::vector<WorkItem> work;
    size_t lds = 0;
    if (prec == 0) {
        std::vector<CoarseArgs<double>> recs;
        for (int s = 0; s < b->S; s++) {
            MG &g = *step_solver(b->sims[s], k);
            if (g.big_exact()) continue;
            work.push_back(WorkItem{(int)recs.size(), 0, s});
            recs.push_back(g.coarse_args<double>(g.L, tol2));
            lds = std::max(lds, g.coarse_lds(0));
        }
        BLaunch L = add_launch(B, recs, work, lds, nullptr);
        L.grid = std::max(1, L.nw);
        return L;
    }
    std::vector<CoarseArgs<float>> recs;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        if (g.big_exact()) continue;
        work.push_back(WorkItem{(int)recs.size(), 0, s});
        recs.push_back(g.coarse_args<float>(g.LF, tol2));
        lds = std::max(lds, g.coarse_lds(1));
    }
    BLaunch L = add_launch(B, recs, work, lds, nullptr);
    L.grid = std::max(1, L.nw);
    return L;
}

// Phase ph of the multi-workgroup exact solve of the members whose coarsest
// level takes it.
template <class T>
static BLaunch x_launch_t(vdi_batch *b, Blob &B, int k, int ph) {
    std::vector<XArgs<T>> recs;
    std::vector<WorkItem> work;
    size_t lds = 0;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        if (!g.big_exact()) continue;
        XArgs<T> a;
        // the throughput layout: the batch keeps every CU busy
        if constexpr (std::is_same<T, double>::value) a = g.x_args<double>(g.L, ph, 0);
        else a = g.x_args<float>(g.LF, ph, 0);
        const int nb = mgk::xsolve_grid(a);
        for (int t = 0; t < nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
        recs.push_back(a);
        lds = std::max(lds, g.x_lds(ph));
    }
    return add_launch(B, recs, work, lds, mgk::xsolve_kernel(sizeof(T) == 4, true));
}

static BLaunch vec_launch(vdi_batch *b, Blob &B, int k, int op, const double *coef_base) {
    std::vector<VecArgs> recs;
    std::vector<WorkItem> work;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        VecArgs a = g.vec_args(op, coef_base ? coef_base + 8 * s + 2 * k : nullptr);
        for (int t = 0; t < a.nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
        recs.push_back(a);
    }
    BLaunch L = add_launch(B, recs, work, 0, nullptr);
    L.grid = std::min(L.grid, 2048);
    return L;
}

// Norm reductions after k_begin (begin = 1) or after a cycle's level-0 up
// kernel (begin = 0; members whose coarse kernel owns level 0 reduce there).
static BLaunch norm_launch(vdi_batch *b, Blob &B, int k, int begin, double tol2) {
    std::vector<NormArgs> recs;
    std::vector<WorkItem> work;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        if (!begin && g.lc == 0) continue;
        recs.push_back(g.norm_args(begin, tol2));
    }
    BLaunch L = add_launch(B, recs, work, 0, nullptr);
    L.nw = (int)recs.size();
    return L;
}

static BLaunch begin_launch(vdi_batch *b, Blob &B, int k, double tol2) {
    std::vector<BeginArgs> recs;
    std::vector<WorkItem> work;
    for (int s = 0; s < b->S; s++) {
        MG &g = *step_solver(b->sims[s], k);
        BeginArgs a = g.begin_args(tol2);
        for (int t = 0; t < a.nb; t++) work.push_back(WorkItem{(int)recs.size(), t, s});
        recs.push_back(a);
    }
    BLaunch L = add_launch(B, recs, work, 0, nullptr);
    L.grid = std::min(L.grid, 2048);
    return L;
}

// Builds every launch of the batched step into one device blob.
static int build_plan(vdi_batch *b) {
    b->drop_graphs();
    Blob B;
    b->recs_host.clear();
    for (int s = 0; s < b->S; s++) {
        vdi_ctx *c = b->sims[s];
        StepRec a = step_record(c, c->P.dt_reg);
        vdi_ctx::Readback *rbd = b->rb_map + s;
        a.rb_st = rbd->st;
        a.rb_gas = rbd->gas;
        b->recs_host.push_back(a);
    }
    b->recs_off = B.put(b->recs_host.data(), b->recs_host.size() * sizeof(StepRec));
    for (int k = 0; k < SK_COUNT; k++) {
        std::vector<WorkItem> work;
        for (int s = 0; s < b->S; s++) {
            const int nt = stepk::tiles(k, b->sims[s]->g);
            for (int t = 0; t < nt; t++) work.push_back(WorkItem{s, t, s});
        }
        BLaunch L;
        L.rec_off = b->recs_off;
        L.work_off = B.put(work.data(), work.size() * sizeof(WorkItem));
        L.nw = (int)work.size();
        size_t lds = 0;
        if (k == SK_GAS)
            for (vdi_ctx *c : b->sims) lds = std::max(lds, c->gas_lds);
        L.lds = lds;
        const bool per_sim = k == SK_GAS || k == SK_VISCOUS_BC || k == SK_GATHER;
        L.grid = per_sim ? b->S : mgk::persistent_grid(stepk::kernel_b(k), lds, L.nw);
        b->step_l[k] = L;
    }
    const double tol2 = b->sims[0]->rel_tol * b->sims[0]->rel_tol;
    const double *coef = b->coef_dev.as<double>();
    for (int k = 0; k < 4; k++) {
        SolvePlan &P = b->solve[k];
        P = SolvePlan{};
        P.used = solver_used(b->sims[0], k);
        if (!P.used) continue;
        apply_tiles(b, k);
        P.fem = k == 3;
        P.mixed = b->form == VDI_FORM_MIXED;
        for (vdi_ctx *c : b->sims) P.maxlc = std::max(P.maxlc, step_solver(c, k)->lc);
        P.extrap = vec_launch(b, B, k, VEC_EXTRAP, coef);
        if (P.fem) P.dirichlet = vec_launch(b, B, k, VEC_DIRICHLET, nullptr);
        P.begin = begin_launch(b, B, k, tol2);
        P.nbegin = norm_launch(b, B, k, 1, tol2);
        P.solve_end = vec_launch(b, B, k, VEC_SOLVE_END, nullptr);
        if (!P.mixed) {
            for (int l = 0; l < P.maxlc; l++) {
                P.down.push_back(level_launch(b, B, k, 0, 0, l, 0, tol2));
                P.up.push_back(level_launch(b, B, k, 0, 1, l, 0, tol2));
            }
            P.coarse = coarse_launch(b, B, k, 0, tol2);
            for (int ph = 1; ph <= 3; ph++) P.xs[ph - 1] = x_launch_t<double>(b, B, k, ph);
            P.ncycle = norm_launch(b, B, k, 0, tol2);
        } else {
            for (int par = 0; par < 2; par++) {
                P.down0[par] = level_launch(b, B, k, 1, 0, 0, par, tol2);
                P.mx[par] = level_launch(b, B, k, 1, 2, 0, par, tol2);
            }
            P.downf.push_back(BLaunch{});
            P.upf.push_back(BLaunch{});
            for (int l = 1; l < P.maxlc; l++) {
                P.downf.push_back(level_launch(b, B, k, 1, 0, l, 0, tol2));
                P.upf.push_back(level_launch(b, B, k, 1, 1, l, 0, tol2));
            }
            P.coarsef = coarse_launch(b, B, k, 1, tol2);
            for (int ph = 1; ph <= 3; ph++) P.xsf[ph - 1] = x_launch_t<float>(b, B, k, ph);
            P.ncycle = norm_launch(b, B, k, 0, tol2);
        }
    }
    VDI_TRY(b->blob.alloc(B.bytes.size()));
    VDI_HIP(hipMemcpy(b->blob.p, B.bytes.data(), B.bytes.size(), hipMemcpyHostToDevice));
    b->base = b->blob.as<char>();
    b->plan_ok = true;
    return VDI_OK;
}

template <class R>
static const R *recs_at(vdi_batch *b, const BLaunch &L) {
    return reinterpret_cast<const R *>(b->base + L.rec_off);
}
static const WorkItem *work_at(vdi_batch *b, const BLaunch &L) {
    return reinterpret_cast<const WorkItem *>(b->base + L.work_off);
}

static BatchAux aux_of(vdi_batch *b) {
    BatchAux a;
    a.ts = nullptr;
    a.busy = b->busy_dev.as<unsigned long long>();
    return a;
}

static int launch_level(vdi_batch *b, const BLaunch &L, int prec, int which, bool fem) {
    if (L.nw == 0) return VDI_OK;
    const WorkItem *w = work_at(b, L);
    if (which == 2) return mgk::mx_b(recs_at<MxArgs>(b, L), w, L.nw, fem, L.lds, L.grid, aux_of(b), b->st);
    if (prec == 0) {
        if (which == 0) return mgk::down_b<double>(recs_at<LvArgs<double>>(b, L), w, L.nw, fem, L.lds, L.grid, aux_of(b), b->st);
        return mgk::up_b<double>(recs_at<LvArgs<double>>(b, L), w, L.nw, fem, L.lds, L.grid, aux_of(b), b->st);
    }
    if (which == 0) return mgk::down_b<float>(recs_at<LvArgs<float>>(b, L), w, L.nw, fem, L.lds, L.grid, aux_of(b), b->st);
    return mgk::up_b<float>(recs_at<LvArgs<float>>(b, L), w, L.nw, fem, L.lds, L.grid, aux_of(b), b->st);
}

static int launch_xsolve(vdi_batch *b, const BLaunch *X, int prec) {
    for (int ph = 0; ph < 3; ph++) {
        const BLaunch &L = X[ph];
        if (L.nw == 0) continue;
        if (prec == 0)
            VDI_TRY(mgk::xsolve_b<double>(recs_at<XArgs<double>>(b, L), work_at(b, L), L.nw, L.lds, L.grid, aux_of(b), b->st));
        else
            VDI_TRY(mgk::xsolve_b<float>(recs_at<XArgs<float>>(b, L), work_at(b, L), L.nw, L.lds, L.grid, aux_of(b), b->st));
    }
    return VDI_OK;
}

static int launch_coarse(vdi_batch *b, const BLaunch &L, int prec) {
    if (L.nw == 0) return VDI_OK;
    if (prec == 0)
        return mgk::coarse_b<double>(recs_at<CoarseArgs<double>>(b, L), work_at(b, L), L.nw, L.lds, L.grid, aux_of(b), b->st);
    return mgk::coarse_b<float>(recs_at<CoarseArgs<float>>(b, L), work_at(b, L), L.nw, L.lds, L.grid, aux_of(b), b->st);
}

static int launch_vec(vdi_batch *b, const BLaunch &L) {
    return mgk::vec_b(recs_at<VecArgs>(b, L), work_at(b, L), L.nw, L.grid, aux_of(b), b->st);
}

static int launch_step(vdi_batch *b, int k) {
    const BLaunch &L = b->step_l[k];
    return stepk::launch_b(k, recs_at<StepRec>(b, L), work_at(b, L), L.nw, L.grid, L.lds, aux_of(b), b->st);
}

// One batched V-cycle of kind k (cycle index c for the mixed ping-pong).
static int launch_cycle(vdi_batch *b, int k, int c) {
    const SolvePlan &P = b->solve[k];
    if (!P.mixed) {
        for (int l = 0; l < P.maxlc; l++) VDI_TRY(launch_level(b, P.down[l], 0, 0, P.fem));
        VDI_TRY(launch_coarse(b, P.coarse, 0));
        VDI_TRY(launch_xsolve(b, P.xs, 0));
        for (int l = P.maxlc - 1; l >= 0; l--) VDI_TRY(launch_level(b, P.up[l], 0, 1, P.fem));
        return mgk::norm_b(recs_at<NormArgs>(b, P.ncycle), P.ncycle.nw, b->st);
    }
    const int par = c & 1;
    VDI_TRY(launch_level(b, P.down0[par], 1, 0, P.fem));
    for (int l = 1; l < P.maxlc; l++) VDI_TRY(launch_level(b, P.downf[l], 1, 0, P.fem));
    VDI_TRY(launch_coarse(b, P.coarsef, 1));
    VDI_TRY(launch_xsolve(b, P.xsf, 1));
    for (int l = P.maxlc - 1; l >= 1; l--) VDI_TRY(launch_level(b, P.upf[l], 1, 1, P.fem));
    VDI_TRY(launch_level(b, P.mx[par], 1, 2, P.fem));
    return mgk::norm_b(recs_at<NormArgs>(b, P.ncycle), P.ncycle.nw, b->st);
}

static int launch_solve(vdi_batch *b, int k, int budget) {
    const SolvePlan &P = b->solve[k];
    VDI_TRY(launch_vec(b, P.extrap));
    if (P.fem) VDI_TRY(launch_vec(b, P.dirichlet));
    VDI_TRY(mgk::begin_b(recs_at<BeginArgs>(b, P.begin), work_at(b, P.begin), P.begin.nw, P.begin.grid, aux_of(b),
                         b->st));
    VDI_TRY(mgk::norm_b(recs_at<NormArgs>(b, P.nbegin), P.nbegin.nw, b->st));
    for (int c = 0; c < budget; c++) VDI_TRY(launch_cycle(b, k, c));
    return launch_vec(b, P.solve_end);
}

__global__ void k_busy_reset(unsigned long long *p, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        p[3 * k] = 0;
        p[3 * k + 1] = ~0ull;
        p[3 * k + 2] = 0;
    }
}

static int capture(vdi_batch *b, vdi_batch::GraphEntry &e) {
    hipGraph_t graph;
    VDI_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
    int rc = VDI_OK;
    do {
        if ((rc = enqueue_fetch(b->coef_map, b->coef_dev.as<double>(), 8 * b->S, b->st)) != VDI_OK) break;
        hipLaunchKernelGGL(k_busy_reset, dim3(cdiv(b->S, 256)), dim3(256), 0, b->st,
                           b->busy_dev.as<unsigned long long>(), b->S);
        if ((rc = launch_step(b, SK_SNAPSHOT)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_GAS)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_GODUNOV)) != VDI_OK) break;
        if ((rc = launch_solve(b, 0, e.budget[0])) != VDI_OK) break;
        if ((rc = launch_step(b, SK_CORRECT)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_SIMPSON)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_VISCOUS_BC)) != VDI_OK) break;
        if (b->implicit) {
            if ((rc = launch_step(b, SK_VISCO_RHS_U)) != VDI_OK) break;
            if ((rc = launch_solve(b, 1, e.budget[1])) != VDI_OK) break;
            if ((rc = launch_step(b, SK_VISCO_RHS_V)) != VDI_OK) break;
            if ((rc = launch_solve(b, 2, e.budget[2])) != VDI_OK) break;
            if ((rc = launch_step(b, SK_VISCO_COPY)) != VDI_OK) break;
        }
        if ((rc = launch_step(b, SK_FEM_RHS)) != VDI_OK) break;
        if ((rc = launch_solve(b, 3, e.budget[3])) != VDI_OK) break;
        if ((rc = launch_step(b, SK_UPDATE)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_VISCOUS_BC)) != VDI_OK) break;
        if ((rc = launch_step(b, SK_GATHER)) != VDI_OK) break;
        // the gather wrote the readbacks into mapped host memory; the clocks follow
        if ((rc = enqueue_fetch(reinterpret_cast<const double *>(b->busy_dev.p),
                                reinterpret_cast<double *>(b->busy_map), 3 * b->S, b->st)) != VDI_OK) break;
    } while (0);
    hipError_t ec = hipStreamEndCapture(b->st, &graph);
    if (rc != VDI_OK) return rc;
    VDI_HIP(ec);
    hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    VDI_HIP(ei);
    return VDI_OK;
}

static bool batch_ready(vdi_batch *b) {
    if (b->primed) return true;
    for (vdi_ctx *c : b->sims)
        for (int k = 0; k < 4; k++) {
            if (!solver_used(c, k)) continue;
            if (step_solver(c, k)->scount < 2 || c->budget[k] < 1) return false;
        }
    b->primed = true;
    return true;
}

static int batch_step_graph(vdi_batch *b) {
    if (!b->plan_ok) VDI_TRY(build_plan(b));
    std::vector<double> t0(b->S);
    for (int s = 0; s < b->S; s++) {
        vdi_ctx *c = b->sims[s];
        const double dt = c->P.dt_reg;
        t0[s] = c->time;
        const double ts[4] = {half_time(t0[s], dt), end_time(t0[s], dt), end_time(t0[s], dt), end_time(t0[s], dt)};
        double *co = b->coef_host + 8 * s;
        for (int k = 0; k < 4; k++) {
            double bb = 1, cc = 0;
            if (solver_used(c, k) && step_solver(c, k)->guess_coeffs(ts[k], &bb, &cc) != 1) { bb = 1; cc = 0; }
            co[2 * k] = bb;
            co[2 * k + 1] = cc;
        }
    }
    // Budget: the most any member needed last step, plus one spare cycle --
    // members that converge earlier skip the rest on their device flags, and
    // a redo on the direct path costs far more than one spare cycle.
    int bud[4] = {0, 0, 0, 0};
    for (vdi_ctx *c : b->sims)
        for (int k = 0; k < 4; k++) bud[k] = std::max(bud[k], c->budget[k]);
    for (int k = 0; k < 4; k++)
        if (bud[k] > 0) bud[k] += 1;
    vdi_batch::GraphEntry *e = nullptr;
    for (auto &x : b->graphs)
        if (std::equal(x.budget, x.budget + 4, bud)) e = &x;
    if (!e) {
        vdi_batch::GraphEntry ne{};
        std::copy(bud, bud + 4, ne.budget);
        VDI_TRY(capture(b, ne));
        b->graphs.push_back(ne);
        e = &b->graphs.back();
    }
    VDI_HIP(hipEventRecord(b->ev0, b->st));
    VDI_HIP(hipGraphLaunch(e->exec, b->st));
    VDI_HIP(hipEventRecord(b->ev1, b->st));
    VDI_HIP(hipStreamSynchronize(b->st));
    float ms = 0;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
    (void)hipGetLastError();
    b->batch_steps++;
    int first_err = VDI_OK;
    for (int s = 0; s < b->S; s++) {
        vdi_ctx *c = b->sims[s];
        const vdi_ctx::Readback &r = b->rb[s];
        c->last_newton = r.gas[1];
        c->last_height_passes = r.gas[2];
        b->status[s] = r.gas[0];
        if (r.gas[0] != VDI_OK) {
            // this member's run ends here (the reference stops with the gas
            // layer's error); the others' steps stand
            if (first_err == VDI_OK) first_err = r.gas[0];
            continue;
        }
        bool ok = true;
        for (int k = 0; k < 4; k++) {
            if (!solver_used(c, k)) continue;
            MG &g = *step_solver(c, k);
            const MGState &stt = r.st[k];
            g.total_cycles += stt.cycles;
            c->redo_base[k] = stt.cycles;
            if (stt.done == 1 || stt.done == 3) {
                g.last_cycles = stt.cycles;
                g.solves++;
                c->record_cycles(k, stt.cycles);
            } else {
                ok = false;
            }
        }
        c->steps++;
        if (ok) {
            c->time = end_time(t0[s], c->P.dt_reg);
            continue;
        }
        b->redo_sims++;
        const int rc = redo_step(c, c->P.dt_reg, t0[s]);
        b->status[s] = rc;
        if (rc == VDI_ERR_HIP) return rc;
        if (rc != VDI_OK && first_err == VDI_OK) first_err = rc;
    }
    return first_err;
}

}  // namespace
}  // namespace vdi

extern "C" {

int vdi_batch_create(const vdi_params *params, int count, int form, vdi_batch **out) {
    if (!params || count < 1 || !out || (form != VDI_FORM_FP64 && form != VDI_FORM_MIXED)) return VDI_ERR_ARG;
    *out = nullptr;
    for (int s = 1; s < count; s++)
        if (params[s].x_sym != params[0].x_sym || params[s].implicit_visc != params[0].implicit_visc) return VDI_ERR_ARG;
    std::unique_ptr<vdi_batch> b(new vdi_batch());
    b->S = count;
    b->status.assign(count, VDI_OK);
    b->form = form;
    b->implicit = params[0].implicit_visc;
    VDI_HIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    for (int s = 0; s < count; s++) {
        vdi_ctx *c = nullptr;
        VDI_TRY(vdi_create(&params[s], &c));
        b->sims.push_back(c);
        c->use_graphs = 0;  // the batch captures the steps
        VDI_TRY(vdi_set_form(c, form));
        // the mixed form needs at least one tiled level: a grid of at most
        // kExactPoints cells is solved exactly at once and cannot take it
        if (form == VDI_FORM_MIXED)
            for (int k = 0; k < 4; k++) {
                if (!solver_used(c, k)) continue;
                if (step_solver(c, k)->lc < 1) return VDI_ERR_ARG;
            }
    }
    for (int k = 0; k < 4; k++) {
        int maxlc = 0;
        for (vdi_ctx *c : b->sims)
            if (solver_used(c, k)) maxlc = std::max(maxlc, step_solver(c, k)->lc);
        for (int pr = 0; pr < 2; pr++)
            for (int w = 0; w < 3; w++) b->tiles[k][pr][w].assign(std::max(1, maxlc), 0);
        // default shapes: the first simulation that has the level decides
        for (int l = 0; l < maxlc; l++)
            for (vdi_ctx *c : b->sims) {
                if (!solver_used(c, k)) continue;
                MG &g = *step_solver(c, k);
                if (l >= g.lc) continue;
                // batched launches walk tiles: a streaming shape of a
                // member's own tuning falls back to the default shape
                for (int pr = 0; pr < 2; pr++)
                    for (int w = 0; w < 2; w++) {
                        const int sh = g.tile[pr][w][l];
                        b->tiles[k][pr][w][l] = kTileShapes[sh].stream ? 0 : sh;
                    }
                b->tiles[k][1][2][0] = kTileShapes[g.tile_mx].stream ? 0 : g.tile_mx;
                break;
            }
    }
    VDI_TRY(b->coef_dev.alloc(8 * sizeof(double) * count));
    VDI_TRY(b->busy_dev.alloc(3 * sizeof(unsigned long long) * count));
    VDI_TRY(mapped_host_alloc((void **)&b->coef_host, (void **)&b->coef_map, 8 * sizeof(double) * count));
    VDI_TRY(mapped_host_alloc((void **)&b->rb, (void **)&b->rb_map, sizeof(vdi_ctx::Readback) * count));
    VDI_TRY(mapped_host_alloc((void **)&b->busy_host, (void **)&b->busy_map, 3 * sizeof(unsigned long long) * count));
    VDI_HIP(hipEventCreate(&b->ev0));
    VDI_HIP(hipEventCreate(&b->ev1));
    int dev = 0, khz = 0;
    VDI_HIP(hipGetDevice(&dev));
    VDI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    b->ticks_per_ms = khz > 0 ? khz : 1e5;
    *out = b.release();
    return VDI_OK;
}

int vdi_batch_destroy(vdi_batch *b) {
    if (!b) return VDI_ERR_ARG;
    delete b;
    return VDI_OK;
}

int vdi_batch_size(vdi_batch *b) { return b ? b->S : -1; }

int vdi_batch_sim(vdi_batch *b, int k, vdi_ctx **out) {
    if (!b || !out || k < 0 || k >= b->S) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(b->st));
    *out = b->sims[k];
    return VDI_OK;
}

int vdi_batch_init_fields(vdi_batch *b) {
    if (!b) return VDI_ERR_ARG;
    for (vdi_ctx *c : b->sims) VDI_TRY(vdi_init_fields(c));
    return VDI_OK;
}

int vdi_batch_step(vdi_batch *b) {
    if (!b) return VDI_ERR_ARG;
    for (int s = 0; s < b->S; s++)
        if (b->status[s] != VDI_OK) return VDI_ERR_ARG;  // an ended member must be restarted first
    if (batch_ready(b)) return batch_step_graph(b);
    // the first steps (time history of the initial guesses) run one by one
    int first_err = VDI_OK;
    for (int s = 0; s < b->S; s++) {
        const int rc = vdi_step_forward(b->sims[s], b->sims[s]->P.dt_reg);
        b->status[s] = rc;
        if (rc == VDI_ERR_HIP) return rc;
        if (rc != VDI_OK && first_err == VDI_OK) first_err = rc;
    }
    return first_err;
}

int vdi_batch_member_status(vdi_batch *b, int *status, int n) {
    if (!b || !status || n < b->S) return VDI_ERR_ARG;
    std::copy(b->status.begin(), b->status.end(), status);
    return VDI_OK;
}

int vdi_batch_restart_member(vdi_batch *b, int k) {
    if (!b || k < 0 || k >= b->S) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(b->st));
    vdi_ctx *c = b->sims[k];
    for (int w = 0; w < 4; w++) {
        if (!solver_used(c, w)) continue;
        MG &g = *step_solver(c, w);
        VDI_TRY(g.reset_history());
        VDI_HIP(hipStreamSynchronize(g.st));
    }
    VDI_TRY(vdi_init_fields(c));
    b->status[k] = VDI_OK;
    b->restarts++;
    return VDI_OK;
}

int vdi_batch_step_times(vdi_batch *b, double *graph_ms, double *busy_ms, int n) {
    if (!b || !graph_ms || (busy_ms && n < b->S)) return VDI_ERR_ARG;
    *graph_ms = b->last_ms;
    if (busy_ms)
        for (int s = 0; s < b->S; s++) busy_ms[s] = (double)b->busy_host[3 * s] / b->ticks_per_ms;
    return VDI_OK;
}

int vdi_batch_member_spans(vdi_batch *b, double *span_ms, int n) {
    if (!b || !span_ms || n < b->S) return VDI_ERR_ARG;
    for (int s = 0; s < b->S; s++) {
        const unsigned long long t0 = b->busy_host[3 * s + 1], t1 = b->busy_host[3 * s + 2];
        span_ms[s] = (t0 == ~0ull || t1 < t0) ? 0.0 : (double)(t1 - t0) / b->ticks_per_ms;
    }
    return VDI_OK;
}

int vdi_batch_counters(vdi_batch *b, long *batched_steps, long *redone_sims) {
    if (!b || !batched_steps || !redone_sims) return VDI_ERR_ARG;
    *batched_steps = b->batch_steps;
    *redone_sims = b->redo_sims;
    return VDI_OK;
}

// Tile shapes of every batched level launch: each candidate is timed on the
// batch's own work list (reps launches after one warm-up) and the fastest
// kept.  Level-0 solutions are saved and restored around the timing.
int vdi_batch_autotune(vdi_batch *b, int reps) {
    if (!b || reps < 1) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(b->st));
    std::vector<std::unique_ptr<DevBuf>> saves;
    for (vdi_ctx *c : b->sims)
        for (int k = 0; k < 4; k++) {
            if (!solver_used(c, k)) continue;
            MG &g = *step_solver(c, k);
            saves.emplace_back(new DevBuf());
            VDI_TRY(saves.back()->alloc(sizeof(double) * g.unknowns()));
            VDI_HIP(hipMemcpy(saves.back()->p, g.z(), sizeof(double) * g.unknowns(), hipMemcpyDeviceToDevice));
        }
    hipEvent_t e0, e1;
    VDI_HIP(hipEventCreate(&e0));
    VDI_HIP(hipEventCreate(&e1));
    const double never = -1.0;
    for (int k = 0; k < 4; k++) {
        if (!solver_used(b->sims[0], k)) continue;
        const bool fem = k == 3;
        int maxlc = (int)b->tiles[k][0][0].size();
        for (int pr = 0; pr < 2; pr++) {
            if (pr == 1 && b->form != VDI_FORM_MIXED) continue;
            if (pr == 0 && b->form == VDI_FORM_MIXED) continue;
            for (int w = 0; w < 3; w++)
                for (int l = 0; l < maxlc; l++) {
                    if (w == 2 && (pr == 0 || l > 0)) continue;
                    if (w == 1 && pr == 1 && l == 0) continue;
                    int best = -1;
                    float best_ms = 0;
                    for (int shape = 0; shape < kNumTileShapes; shape++) {
                        if (kTileShapes[shape].stream) continue;  // batched launches walk tiles
                        // the shape must fit every simulation that has the level
                        bool fits = true;
                        for (vdi_ctx *c : b->sims) {
                            MG &g = *step_solver(c, k);
                            if (w < 2 && l >= g.lc) continue;
                            if (g.set_tile(pr, w, w == 2 ? 0 : l, shape) != VDI_OK) fits = false;
                        }
                        if (!fits) continue;
                        for (vdi_ctx *c : b->sims) {
                            MGState zero{};
                            VDI_HIP(hipMemcpy(step_solver(c, k)->state(), &zero, sizeof(MGState), hipMemcpyHostToDevice));
                        }
                        Blob B;
                        BLaunch L = level_launch(b, B, k, pr, w, l, 0, never);
                        DevBuf tmp;
                        VDI_TRY(tmp.alloc(B.bytes.size()));
                        VDI_HIP(hipMemcpy(tmp.p, B.bytes.data(), B.bytes.size(), hipMemcpyHostToDevice));
                        b->base = tmp.as<char>();
                        int rc = launch_level(b, L, pr, w, fem);
                        VDI_HIP(hipEventRecord(e0, b->st));
                        for (int r = 0; r < reps && rc == VDI_OK; r++) rc = launch_level(b, L, pr, w, fem);
                        VDI_HIP(hipEventRecord(e1, b->st));
                        VDI_HIP(hipEventSynchronize(e1));
                        b->base = nullptr;
                        if (rc != VDI_OK) return rc;
                        float ms = 0;
                        VDI_HIP(hipEventElapsedTime(&ms, e0, e1));
                        if (best < 0 || ms < best_ms) { best = shape; best_ms = ms; }
                    }
                    if (best >= 0) {
                        b->tiles[k][pr][w][l] = best;
                        for (vdi_ctx *c : b->sims) {
                            MG &g = *step_solver(c, k);
                            if (w < 2 && l >= g.lc) continue;
                            (void)g.set_tile(pr, w, w == 2 ? 0 : l, best);
                        }
                    }
                }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    size_t q = 0;
    for (vdi_ctx *c : b->sims)
        for (int k = 0; k < 4; k++) {
            if (!solver_used(c, k)) continue;
            MG &g = *step_solver(c, k);
            VDI_HIP(hipMemcpy(g.z(), saves[q++]->p, sizeof(double) * g.unknowns(), hipMemcpyDeviceToDevice));
            MGState zero{};
            VDI_HIP(hipMemcpy(g.state(), &zero, sizeof(MGState), hipMemcpyHostToDevice));
        }
    b->plan_ok = false;
    return VDI_OK;
}

int vdi_batch_tiles(vdi_batch *b, int *out, int n) {
    if (!b || !out) return VDI_ERR_ARG;
    int q = 0;
    for (int k = 0; k < 4; k++)
        for (int pr = 0; pr < 2; pr++)
            for (int w = 0; w < 3; w++)
                for (int v : b->tiles[k][pr][w]) {
                    if (q < n) out[q] = v;
                    q++;
                }
    return q;
}

}  // extern "C"
