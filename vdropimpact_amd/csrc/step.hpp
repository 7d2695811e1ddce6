// step.hpp -- state of one drop-impact simulation on the device, the
// argument record of the step kernels and their launchers (shared by the
// single-simulation context in step.hip and the batch in batch.hip).
#pragma once

#include <vector>

#include "mg.hpp"

namespace vdi {

// Grid constants (fluid_2d.hh:59-80).
struct Geo {
    int m, n, ml, me, ne, x_sym;
    double dx, dy, xsp, ysp, xxsp, yysp;
    double ax, bx, ay, by;
    double rho, rhoinv, mu, muinv;
};

// Device pointers of the cell fields (fields.hh:7-19, one array each).
struct Fields {
    double *u, *v, *p;
    double *ux, *uy, *vx, *vy;
    double *ud, *vd, *ul, *vl, *ur, *vr, *uu, *vu;
    double *c0, *c1, *c2, *c3;
    double *src;   // right-hand side of the MAC, FEM and u-velocity solves
    double *src2;  // right-hand side of the v-velocity solve (m x n)
};

// Gas-layer arrays (gas_layer.hh:26-34, gas_layer_solve.hh:163-173).
struct Gas {
    double *p, *tau, *p_old, *tau_old, *h, *hx, *ht, *fbd, *pg, *pg_old;
    int *status;  // [0] = status code, [1] = Newton steps, [2] = height passes
};

struct GasConst {
    double alpha, mug, Pamb, h0, R, V, sigma, gxsp, gxxsp;
};

// Everything a step kernel needs about one simulation for one step.
struct StepRec {
    Geo g;
    Fields F;
    Gas G;
    GasConst K;
    double dt, viscfac, macfac, prefac, vfx, vfy, rdt, fm, Vframe;
    const double *zmac, *zvx, *zvy, *zfem;
    double *snap;   // snapshot: u, v, p (3 NP) then the gas arrays (ng)
    double *gasbase;
    long np, ng;
    const MGState *mgst[4];  // MAC, VX, VY, FEM (null when absent)
    MGState *rb_st;          // readback: four states, then two gas ints
    int *rb_gas;
    // Godunov leaves the gas-pressure term of the bottom MAC source row to
    // k_mac_bottom, so that it need not wait for the gas update.
    int defer_sbc;
};

// Tiles of the step kernels: Godunov 32 x 8 cells (256 threads, 1D);
// element-wise kernels 64 x 4 threads.
constexpr int kGX = 32, kGY = 8;
constexpr int kEX = 64, kEY = 4;
constexpr int kGasThreads = 1024;
// Default V(nu1, nu2) of the step's solvers [MAC, FEM, VX, VY], measured at
// the 1024 x 512 single-drop workload (steps/s with every solve to 1e-10):
// V(2,2) everywhere 922, V(3,3) 1044-1054, these 1104-1116.  A level launch
// runs all its sweeps in LDS, so a sweep more costs far less than the
// V-cycle it saves; the Helmholtz solves then converge in one or two cycles.
// The last two entries are the sweeps on the restricted levels: V(2,2)
// there makes a nodal cycle 8 % faster at nearly the same convergence
// factor, V(0,1) a Helmholtz cycle about a fifth faster at the same cycle
// count; the cell-centred pressure solve converges too slowly with fewer.
constexpr int kStepNu[4][4] = {{3, 2, 3, 2}, {3, 2, 2, 2}, {4, 2, 0, 1}, {4, 2, 0, 1}};
constexpr int kGasRows = 5;  // rows per thread: me <= 5120
// Height update by Jacobi passes (m <= kGasJacobiM: six LDS arrays of m
// doubles); at most kGasJacobiPasses passes before the serial sweep.
constexpr int kGasJacobiM = 2560;
constexpr int kGasJacobiPasses = 96;

// Kernel identifiers for the batched launcher.
enum StepKernel {
    SK_GODUNOV = 0,
    SK_CORRECT,
    SK_SIMPSON,
    SK_VISCOUS_BC,
    SK_VISCO_RHS_U,
    SK_VISCO_RHS_V,
    SK_VISCO_COPY,
    SK_FEM_RHS,
    SK_UPDATE,
    SK_GAS,
    SK_SNAPSHOT,
    SK_RESTORE,
    SK_GATHER,
    SK_COUNT
};

struct LearnedGuess;
// Learned initial guess of the MAC solve (nn.hip): pairs recorded around a
// direct solve, and the network's guess in place of the extrapolation.
int guess_capture_pre(LearnedGuess &L, MG &g, hipStream_t st);
int guess_after_solve(LearnedGuess &L, MG &g, hipStream_t st);
int guess_infer(LearnedGuess &L, MG &g, hipStream_t st);
int guess_mode(const LearnedGuess *L);
void guess_destroy(LearnedGuess *L);

namespace stepk {
// Work items (record, tile) of kernel k for one simulation.
int tiles(int k, const Geo &g);
// Batched launch of kernel k over the work list.
int launch_b(int k, const StepRec *recs, const WorkItem *w, int nw, int grid, size_t lds, BatchAux aux,
             hipStream_t st);
int threads(int k);
const void *kernel_b(int k);
size_t gas_lds(const Geo &g);
}  // namespace stepk

}  // namespace vdi

// One simulation: device state, the four solvers and the captured steps.
struct vdi_ctx {
    vdi_params P{};
    vdi::Geo g{};
    vdi::GasConst K{};
    hipStream_t st = nullptr;
    // Captured steps run the gas update beside the Godunov passes, and the
    // two viscous solves beside each other, as parallel graph branches: the
    // second branch is recorded on st2 between a fork and a join event.
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_gfork = nullptr, ev_gjoin = nullptr;
    hipEvent_t ev_god = nullptr, ev_bcjoin = nullptr;
    int par_visco = 1;
    double time = 0, Vframe = 0;
    double rel_tol = 1e-10;
    int max_cycles = 200;
    int form = VDI_FORM_FP64;
    long steps = 0;
    int last_newton = 0;
    int last_height_passes = 0;
    vdi::DevBuf fields;  // 19 padded arrays
    vdi::DevBuf srcbuf, gasbuf, zbuf, statusbuf;
    vdi::Fields F{};
    vdi::Gas G{};
    double *zmac = nullptr, *zfem = nullptr, *zvx = nullptr, *zvy = nullptr;
    vdi::MG mac, fem, vx, vy;
    int *host_status = nullptr;
    size_t gas_lds = 0;
    vdi::KernelTimer timer;
    // Captured steps (one graph per cycle budget, time step and form).
    struct GraphEntry {
        int budget[4];
        double dt;
        int form;
        hipGraphExec_t exec;
        int slot_a, slot_b;  // timer slots owned by the graph
        // couplings and coarse inverses of the two viscous solves the graph's
        // kernels were recorded with (re-adopted before a replay)
        double vfx, vfy;
        std::shared_ptr<vdi::CoarseFactors> vinv[2];
        long used;  // replay clock of the last launch (least recently used is evicted)
    };
    // Cached graphs (distinct budget tuples, time steps and forms); the least
    // recently used one is destroyed when a new one would exceed this.
    static constexpr int kMaxGraphs = 8;
    long graph_clock = 0;
    struct Readback {
        vdi::MGState st[4];
        int gas[3];
        int pad;
    };
    int use_graphs = 1;
    int budget[4] = {0, 0, 0, 0};
    // cycles each solve ran in a captured step that is being redone: the
    // redo's solve starts from those iterates, so its budget becomes these
    // plus the cycles the redo needed
    int redo_base[4] = {0, 0, 0, 0};
    // Cycles of the last kHist solves of each system; a captured step's
    // budget is their maximum.  A count that alternates (3, 4, 3, ...) would
    // otherwise send every rise through a redo, which costs a whole step on
    // the direct path, while a cycle too many costs only its launches
    // (their kernels return at once on the converged flag).
    static constexpr int kHist = 16;
    int hist[4][kHist] = {};
    int hist_n[4] = {0, 0, 0, 0};
    void record_cycles(int k, int cycles) {
        hist[k][hist_n[k] % kHist] = cycles;
        hist_n[k]++;
        int b = 1;
        for (int q = 0; q < std::min(hist_n[k], kHist); q++) b = std::max(b, hist[k][q]);
        budget[k] = b;
    }
    std::vector<GraphEntry> graphs;
    vdi::DevBuf snap, coef_dev;
    // Host memory the captured graph's kernels read and write directly
    // (pinned, mapped, coherent): the extrapolation coefficients of the next
    // step (2 per solver), fetched into coef_dev by the graph's first kernel,
    // and the step's readback, written by its last kernel.  A captured step
    // thereby holds kernel nodes only.
    double *coef_host = nullptr, *coef_map = nullptr;
    Readback *rb = nullptr, *rb_map = nullptr;
    long graph_steps = 0, redo_steps = 0;
    vdi::LearnedGuess *guess = nullptr;  // optional learned MAC initial guess
    vdi::DevBuf tracers;                  // 2 ntrace positions (fluid_2d::tm)
    int ntrace = 0;
    void drop_graphs() {
        for (auto &e : graphs) (void)hipGraphExecDestroy(e.exec);
        graphs.clear();
        timer.next = timer.reserved = 0;
    }
    ~vdi_ctx() {
        if (st) (void)hipStreamSynchronize(st);
        drop_graphs();
        if (host_status) (void)hipHostFree(host_status);
        if (coef_host) (void)hipHostFree(coef_host);
        if (rb) (void)hipHostFree(rb);
        if (guess) vdi::guess_destroy(guess);
        if (st2) (void)hipStreamSynchronize(st2);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_gfork) (void)hipEventDestroy(ev_gfork);
        if (ev_gjoin) (void)hipEventDestroy(ev_gjoin);
        if (ev_god) (void)hipEventDestroy(ev_god);
        if (ev_bcjoin) (void)hipEventDestroy(ev_bcjoin);
        if (st2) (void)hipStreamDestroy(st2);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace vdi {
// Record of one step of c with time step dt (step.hip).
StepRec step_record(vdi_ctx *c, double dt);
// Solves in step order: MAC, VX, VY, FEM.
MG *step_solver(vdi_ctx *c, int k);
bool solver_used(vdi_ctx *c, int k);
// Time of the half and full step (two half steps, fluid_2d.cc:102, :251).
double half_time(double t0, double dt);
double end_time(double t0, double dt);
// Restores the snapshot and runs the step on the direct path (a redo).
int redo_step(vdi_ctx *c, double dt, double t0);
// visco_impl.cc:26-30
int visco_setup(MG &g, double vfx, double vfy);
// Pinned host memory that kernels read and write directly (mapped, coherent):
// *host for the host, *dev for kernel arguments.
int mapped_host_alloc(void **host, void **dev, size_t bytes);
// dst[0..n) = src[0..n) of doubles (or 8-byte words), one capturable kernel.
int enqueue_fetch(const double *src, double *dst, int n, hipStream_t st);
}  // namespace vdi
