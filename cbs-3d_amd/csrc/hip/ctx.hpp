// Internal definitions shared by the translation units of libcbs3d_hip.so.
//
// Device data layout (E elements, F fluid elements, N nodes):
//  - per-element arrays in element order: connectivity as int4 (0-based
//    nodes), dNkdx (12), detJ, annxf (16), elcoe_e (4), fedge packed in one
//    uint32 (one byte per local face), materials, pdiagE (4), delte;
//  - the pressure apply streams fluid elements in colour-major "slot" order
//    (the reference's Coloring order with solids removed): connectivity and
//    the six coefficient planes gstif[p*F + k] are read fully coalesced;
//  - node-incidence lists sorted by element (packed e*4+a) drive every
//    element-order assembly as a per-node gather, so each node sums its
//    element contributions in exactly the reference's serial order;
//  - node vectors are plain SoA arrays; velocity is (3,N) interleaved like
//    the reference's unkno.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cbs_hip.h"

namespace cbsdev
{
    constexpr int kBlock = 256;          // threads per block for node/element kernels
    constexpr int kReduceBlocks = 2048;  // grid of the grid-stride reduction kernels
    constexpr int kSlotWindow = 512;     // elements (threads) per block of the line-group element passes
    // meshes from this many elements take the line-group element passes
    // (measured: 2.3% faster multigrid steps at 24.6 M tets; 1-2% slower
    // steps at 48 K-786 K tets, where the passes are latency-bound)
    constexpr int kLineGroupMinElems = 1 << 22;

    // Stages for error precedence (earlier stage, loop, index, check wins).
    enum Stage : uint32_t
    {
        ST_TIMESTEP = 1, ST_LHS = 2, ST_STEP1 = 3, ST_STEP2RHS = 4, ST_CG = 5,
        ST_STEP3 = 6, ST_STEP4 = 7, ST_CONV = 8, ST_PTERMS = 9
    };

    // Device-resident scalars of a context.
    struct DevScalars
    {
        unsigned long long err;  // first error key, ~0ull when none
        double dtreal;
        double dt_min;
        double hb[15];
        // pressure CG
        double rz_old, alpha, beta, pAp;
        double initial_l2, final_l2, rel_l2, max_abs;
        int iter, done, converged, max_iter;
        int conv_test, pad0;
        double rel_tol, abs_tol;
        // pending preconditioner error of the current iteration (raised only
        // when the iteration does not converge)
        unsigned long long pending_err;
        // mixed-precision inner solve scalars
        // f_done: paused (the recursive residual has dropped by f_tol since
        // the last true residual) or finished; f_refresh: this segment's
        // true residual is being recomputed (the others are left alone)
        double f_rz_old, f_alpha, f_beta, f_initial, f_final, f_rel, f_tol;
        int f_iter, f_done, f_max_iter, f_refresh;
        // refinement
        double true_l2, true_max;
        // deferred iterate update: the previous iteration's alpha p is not
        // yet in x (it is added with the current one, or by a flush)
        int x_pending, pad2;
        double alpha_prev;
    };

    // Contiguous item ranges [beg[b], end[b]) of one segment each: the block
    // decomposition of the reductions over nodes, elements or shared nodes.
    // Every segment is decomposed exactly as a context holding that segment
    // alone would be, so per-segment reductions are the same sums.
    struct BlockMap
    {
        const int32_t* beg;
        const int32_t* end;
        const int32_t* seg;
    };

    __host__ __device__ inline unsigned long long err_key(uint32_t stage, uint32_t loop, uint32_t index,
                                                          uint32_t check)
    {
        return (static_cast<unsigned long long>(stage) << 48) | (static_cast<unsigned long long>(loop) << 40) |
               (static_cast<unsigned long long>(index) << 8) | static_cast<unsigned long long>(check);
    }
}

struct cbs_hip_blocks
{
    int32_t n = 0;                       // blocks
    int32_t *beg = nullptr, *end = nullptr, *seg = nullptr;
    int32_t* seg_ptr = nullptr;          // device: nseg+1 block offsets per segment
    std::vector<int32_t> h_seg_ptr;
    cbsdev::BlockMap map() const { return {beg, end, seg}; }
};

// The assembled pressure operator of the PCG apply (k_sell.hip): sliced
// ELLPACK over row tiles whose x is staged in LDS, on vectors in solver
// order (gid[s] is the node of solver index s; every row tile's rows are
// consecutive and each segment keeps its node range).  Row tile t has rows
// row_base[t] .. row_base[t] + row_cnt[t] and stages them, then its halo
// halo[halo_ptr[t] .. halo_ptr[t+1]) (solver indices); it holds slices
// slice_ptr[t] .. slice_ptr[t+1]; slice s covers entries slice_off[s] ..
// slice_off[s+1] (64 rows times an even entry count).  Values are rebuilt
// at every solve: entry q sums gstif_t at cidx[cptr[q] .. cptr[q+1]), or,
// in a segment whose elements share one time step, is that step times the
// same sum of gstifE_t (base64).
struct cbs_hip_sell
{
    int32_t ntiles = 0, lmax = 0, max_slices = 0;
    int64_t nnz = 0, entries = 0, nstage = 0, ncontrib = 0;
    int32_t *row_base = nullptr, *row_cnt = nullptr, *halo_ptr = nullptr, *halo = nullptr, *slice_ptr = nullptr;
    int32_t* gid = nullptr;
    int64_t* slice_off = nullptr;
    uint32_t *cols = nullptr, *cptr = nullptr, *cidx = nullptr;
    double* v64 = nullptr;
    float* v32 = nullptr;
    // Values per unit time step: entry q sums gstifE_t at the same
    // contributions (rebuilt when the element pressure terms are); a segment
    // whose elements all have one time step gets that step times them.
    double* base64 = nullptr;
    bool base_stale = true;
    int64_t* seg_entry = nullptr;    // nseg+1 entry offsets per segment
    int32_t* seg_elem0 = nullptr;    // first element of each segment
    int32_t* dt_uniform = nullptr;   // per segment, set by the values refresh
    std::vector<int32_t> h_seg_tile;  // nseg+1 row-tile offsets per segment
    int32_t* seg_tile_ptr = nullptr;
    int32_t* tile_seg = nullptr;      // only when nseg > 1
    // launch shape per precision (0 FP64, 1 FP32), chosen by the apply
    // autotuner at the first solve: threads per tile, column pairs per load batch
    struct Config
    {
        int32_t threads;
        int32_t batch;
        bool tuned;
        float tune_ms[6];
        // the PCG iteration's direction update: 1 fused into the apply's
        // staging, 0 a kernel of its own; -1 not yet timed.  fuse_pin >= 0
        // fixes it (cbs_hip_set_direction_fusion).  fuse_ms: the timed
        // (separate, fused) milliseconds per iteration's apply and direction.
        int32_t fuse = -1;
        int32_t fuse_pin = -1;
        float fuse_ms[2] = {0.0f, 0.0f};
    };
    Config cfg[2] = {{512, 8, false, {}}, {512, 8, false, {}}};
};

// The row tiles of the multi-column product (k_sell.hip, built at the first
// multi-RHS solve): each row tile of the assembled operator cut into parts of
// consecutive slices whose staged set fits node_bytes per staged node in LDS.
// A part keeps its tile's slices and values (slice_ptr indexes the
// operator's slices) and has its own rows, halo and local column ids.
struct cbs_hip_sell_parts
{
    int32_t nparts = 0, lmax = 0, max_slices = 0, node_bytes = 0;
    int32_t *row_base = nullptr, *row_cnt = nullptr, *halo_ptr = nullptr, *halo = nullptr, *slice_ptr = nullptr;
    uint32_t* cols = nullptr;
};

// One output snapshot slot: the solution arrays a VTU file holds, copied
// within device memory on the compute stream, then to pinned host memory on
// the copy stream.  Layout of d and h: pres (N), temperature (N), unkno (3N),
// velocity magnitude (N).
struct cbs_snapshot
{
    cbs_hip_ctx* ctx = nullptr;
    double* d = nullptr;
    double* h = nullptr;
    hipEvent_t copied = nullptr, ready = nullptr;
    bool held = false;
};

struct cbs_hip_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    cbs_config cfg{};
    int32_t N = 0, E = 0, B = 0, F = 0, ncolor = 0, bc_fixed = 0, npoin_wall = 0;
    std::string err;

    // Segments.  A context built by cbs_hip_create is one segment; a batch
    // context (cbs_hip_create_batch) holds the disjoint union of several
    // cases, one segment each: nodes [h_seg_node[s], h_seg_node[s+1]) and
    // elements [h_seg_elem[s], h_seg_elem[s+1]) belong to member s.  Scalars
    // the reference computes globally (minimum time step, dtreal, norms, PCG
    // scalars) are per segment; scal points at nseg DevScalars.
    int32_t nseg = 1;
    std::vector<int32_t> h_seg_node, h_seg_elem, h_seg_tile;
    std::vector<double> h_seg_ani;
    uint16_t* seg_of_node = nullptr;  // only when nseg > 1
    uint16_t* seg_of_elem = nullptr;
    double* seg_ani = nullptr;
    cbs_hip_blocks nblk, eblk;         // node and element block maps
    std::vector<cbs_cg_result> seg_result;  // last solve, per segment
    std::vector<double> seg_rtime;

    // element order
    int4* conn = nullptr;
    int32_t* mat = nullptr;
    double *rho = nullptr, *mu = nullptr, *kcond = nullptr, *rcp = nullptr, *alpha = nullptr, *qvol = nullptr;
    double *dNkdx = nullptr, *detJ = nullptr, *annxf = nullptr, *elcoe = nullptr, *alen_e = nullptr;
    uint32_t* fedge = nullptr;
    double *pdiagE = nullptr, *gstifE = nullptr, *delte = nullptr;

    // colour-major fluid slots
    std::vector<int32_t> h_fcolor_ptr;  // ncolor+1 offsets into the slot arrays
    int32_t* slot_elem = nullptr;       // element (0-based) of each slot
    int32_t* elem_slot = nullptr;       // slot of each element, -1 for solids
    int4* slot_conn = nullptr;
    double* gstifE_s = nullptr;         // 6 planes of F
    double* gstif_s = nullptr;          // 6 planes of F
    bool gstif_slots_stale = false;     // gstif_s older than the LHS update
    bool gstif_tiles_stale = false;     // gstif_t older than the LHS update
    // gstif set by cbs_hip_set_field: the operator's values are summed from
    // gstif_t (not dt times the gstifE sums) until the next LHS update
    bool gstif_manual = false;

    // element tiles: fluid elements grouped into spatially compact tiles by
    // recursive coordinate bisection; inside a tile the elements are ordered
    // by (colour, element) and address the tile's nodes through 16-bit local
    // ids.  A tile's nodes are listed interior first (nodes no other tile
    // references), then the nodes it shares with other tiles, each group in
    // ascending node id.  The exact apply (k_apply_tiled) gives shared nodes
    // one value per colour in the side buffer sb[c*nshared + s], summed in
    // colour order afterwards, so every node sees the reference's sequence
    // of additions.  The tiles also define the row tiles of the assembled
    // operator (sell) and the tile-slot order of gstif_t.
    int32_t ntiles = 0, tile_lmax = 0, nshared = 0, tile_threads = 1024;
    int32_t *tnode_ptr = nullptr, *tnodes = nullptr, *tsidx = nullptr, *tcol_ptr = nullptr;
    int32_t *tslot_elem = nullptr, *elem_tslot = nullptr, *snode = nullptr;
    ushort4* lidx = nullptr;
    double *gstifE_t = nullptr, *gstif_t = nullptr, *sb = nullptr;
    double* dotpart = nullptr;  // one per row tile: partial dots of the fused apply
    int32_t ndotpart = 0;
    cbs_hip_sell sell;                      // PCG operator (assembled)
    cbs_hip_sell_parts sell_parts;          // its row tiles for the multi-column product

    // Node-sum slots of the two-pass assemblies.  Element passes write the
    // term of (element e, vertex a) to slot pos[e*4 + a] of every plane;
    // node passes read them back coalesced: nodes are cut into slices of 64,
    // and slot k*64 + l of slice s (from off[s]; (off[s+1] - off[s]) / 64
    // columns) holds the k-th term of node 64 s + l, in the order the
    // reference adds it.  Slots past a node's last term stay +0.0.
    //  - slots_e: every incidence in element order (LHS, Step 2, Step 3);
    //  - slots_c: fluid incidences in (colour, element) order (Step 1).
    struct NodeSlots
    {
        int64_t nslots = 0;
        int64_t* off = nullptr;   // nslice + 1
        uint32_t* pos = nullptr;  // 4 E
        double* q = nullptr;      // 3 planes of nslots
        // The element passes' stores as whole lines: block b of kSlotWindow
        // elements writes line groups grp_ptr[b] .. grp_ptr[b+1]; group g
        // is the 16 slots from 16*grp_line[g], lane j taking the block's
        // incidence grp_src[16g + j] (local element * 4 + vertex; 0xFFFF:
        // a slot no incidence of the block writes).  Null: not built.
        int32_t* grp_ptr = nullptr;
        uint32_t* grp_line = nullptr;
        uint16_t* grp_src = nullptr;
    };
    NodeSlots slots_e, slots_c;
    int32_t* elem_fluid = nullptr;  // fluid elements ascending

    // node-incidence lists in element order
    int32_t* inc_ptr = nullptr;  // N+1
    int32_t* inc = nullptr;      // 4E, packed e*4+a

    // node flags: bit0 touches fluid, bit1 touches solid, bit2 fixed pressure
    uint8_t* nflag = nullptr;

    // nodal state
    double *unkno = nullptr, *unkn1 = nullptr, *pres = nullptr, *pres1 = nullptr;
    double *temp = nullptr, *temp1 = nullptr, *rhs = nullptr, *rhs1 = nullptr;
    double *deltp = nullptr, *deltp1 = nullptr, *deltp2 = nullptr, *beta1 = nullptr, *vvis = nullptr;
    double *velcp = nullptr, *alen = nullptr, *unit_vec = nullptr;
    // nodal lengths from the geometry (compute_nodal_lengths_from_elements,
    // TimeStep.cpp:319-362), computed once at create; the field alen takes
    // them at every time step that computes them (not the fixed-dt branch)
    double* alen_static = nullptr;
    double *elcoe2 = nullptr, *elcoe2p = nullptr, *pdiag = nullptr, *velocity = nullptr;

    // pressure CG work, in solver order (cbs_hip_sell): the iterate, the
    // right-hand side, pdiag and the node flags of the solve
    double *sv_x = nullptr, *sv_b = nullptr, *sv_pdiag = nullptr;
    uint8_t* sv_flag = nullptr;
    double *cg_r = nullptr, *cg_p = nullptr, *cg_ap = nullptr, *tmp_x = nullptr;
    double* cg_p2 = nullptr;  // direction of odd iterations (FP64 form)
    // mixed form: FP32 recursive residual, direction, Ap, inverse Jacobi
    // diagonal and pdiag (solver order)
    float *f_r = nullptr, *f_p = nullptr, *f_ap = nullptr, *f_dinv = nullptr, *f_z = nullptr;
    float* f_p2 = nullptr;    // direction of odd iterations (mixed form)
    double* partial = nullptr;    // 16 per block of nblk / eblk: block partials of the reductions
    // non-free pressure nodes: (node, mode) with mode 0 inactive, 1 fixed
    int32_t nnonfree = 0;
    int32_t* nf_node = nullptr;
    int32_t* nf_mode = nullptr;
    double* nf_val = nullptr;
    // fixed pressure rows (Boundary::applyPressure / identity rows), last writer per node
    int32_t npres_bc = 0;
    int32_t* pbc_node = nullptr;
    double* pbc_val = nullptr;

    // forward-model extensions of the adjoint workload (off unless set):
    // a nodal body force added to Step 1 (acceleration, per node and
    // component) and Fourier coefficients added to the BC-508 inflow profile
    double* bforce = nullptr;      // 3N, null when off
    double* bforce_buf = nullptr;  // the force's device buffer (kept when switched off)
    bool constraint_zero = false;  // adjoint solves: non-free nodes held at 0
    double* node_mass = nullptr;   // N lumped fluid mass (sum of elcoe over fluid elements)
    std::vector<double> h_bcv_val;        // prescribed values as built (3 per BC slot)
    std::vector<int32_t> h_inflow_slot;   // BC slots whose final value is the 508 profile
    std::vector<double> h_inflow_yz;      // (y, z) of those slots' nodes
    std::vector<double> h_inflow_coef;    // 16 when set
    struct AdjointState* adj = nullptr;   // checkpoints, tapes and adjoint vectors
    struct AmgState* amg = nullptr;       // multigrid hierarchy (built at the first multigrid solve)

    // boundary programmes
    int32_t nbcv = 0;
    int32_t *bcv_node = nullptr, *bcv_flags = nullptr, *bcv_sym_ptr = nullptr, *bcv_out_ptr = nullptr;
    double *bcv_val = nullptr, *bcv_sym_n = nullptr, *bcv_out_n = nullptr;
    int32_t ntbc = 0;
    int32_t* tbc_node = nullptr;
    double* tbc_val = nullptr;
    int32_t nhf = 0;  // nodes receiving BC-532 heat flux
    int32_t *hf_slot = nullptr, *hf_ptr = nullptr;  // hf_slot: node -> row of hf_ptr, or -1
    double* hf_val = nullptr;

    cbsdev::DevScalars* scal = nullptr;    // device, nseg entries
    cbsdev::DevScalars* h_scal = nullptr;  // pinned host mirror, nseg entries

    // wall-clock seconds of the phases of the context's construction
    std::vector<double> create_seconds;

    // output snapshots (cbs_hip_snapshot_*): two slots, a copy stream
    hipStream_t copy_stream = nullptr;
    cbs_snapshot snap[2];
    std::mutex snap_mu;
    std::condition_variable snap_cv;

    // last solve bookkeeping (host)
    int32_t last_cg_iterations = 0;
    double last_cg_initial_l2 = 0, last_cg_final_l2 = 0, last_cg_relative_l2 = 0, last_cg_max_abs = 0;
    double rtime = 0.0;

    std::vector<void*> allocations;
    size_t device_bytes = 0;

    // apply timing (HIP events on `stream`), see cbs_hip_set_profiling
    bool profile = false;
    std::vector<hipEvent_t> events;  // pairs: begin, end
    int events_used = 0;
    // set by timer_begin: the next apply launch takes the pair as the
    // kernel's own start and end events (hipExtLaunchKernelGGL)
    bool timer_pending = false;
    double apply_ms[2] = {0.0, 0.0};
    int64_t apply_n[2] = {0, 0};
    std::vector<int> event_prec;     // precision of each recorded pair
    std::vector<int> event_idx;      // position of the timed apply in its chunk of iterations
    // phase timing of pressure solves (profiling on): each mark records an
    // event closing the kernel group since the previous mark, labelled with
    // the ConjugateGradient::Result phase of that group (see PhaseMark)
    std::vector<hipEvent_t> ph_events;
    std::vector<int> ph_label, ph_idx;
    int ph_used = 0;
    double ph_sampled_ms[5] = {0, 0, 0, 0, 0};  // per phase, over the sampled iterations
    double ph_direct_ms[5] = {0, 0, 0, 0, 0};   // per phase, groups timed every time
    int ph_sampled_iters = 0;

    // PCG iterations launched as captured HIP graphs (cbs_hip_set_graphs):
    // per solver form (0 FP64, 1 mixed, 2 multigrid) one block of
    // iterations captured after the apply autotuner has fixed the launch
    // shape.
    struct PcgGraph
    {
        hipGraphExec_t exec = nullptr;
    };
    PcgGraph pcg_graph[3];
    bool use_graphs = true;
};

namespace cbsdev
{
    // launchers implemented in k_steps.hip
    void launch_build_pressure_terms(cbs_hip_ctx* c);
    void launch_static_alen(cbs_hip_ctx* c);
    void launch_timestep(cbs_hip_ctx* c, int iitime);
    void launch_lhs(cbs_hip_ctx* c);
    void launch_step1_rhs(cbs_hip_ctx* c);
    void launch_step1_update(cbs_hip_ctx* c);
    void launch_step2_rhs(cbs_hip_ctx* c);
    void launch_step3(cbs_hip_ctx* c);
    void launch_step4_rhs(cbs_hip_ctx* c);
    void launch_step4_update(cbs_hip_ctx* c);
    void launch_bc(cbs_hip_ctx* c, int which);
    void launch_velocity_magnitude(cbs_hip_ctx* c);
    void launch_convergence(cbs_hip_ctx* c);
    void launch_fill(cbs_hip_ctx* c, double* p, int64_t n, double v);
    void launch_copy(cbs_hip_ctx* c, double* dst, const double* src, int64_t n);

    // launchers implemented in k_pressure.hip
    void launch_apply(cbs_hip_ctx* c, const double* x, double* acc, const int* done_flag);
    void launch_apply_fast(cbs_hip_ctx* c, const double* x, double* acc);
    void configure_tiled_apply(cbs_hip_ctx* c);
    void launch_apply_finish(cbs_hip_ctx* c, const double* x, double* y);
    int solve_pressure_fp64(cbs_hip_ctx* c, cbs_cg_result* res);
    int solve_pressure_mixed(cbs_hip_ctx* c, const cbs_cg_options* o, cbs_cg_result* res);
    void launch_gstif(cbs_hip_ctx* c);
    void ensure_gstif_slots(cbs_hip_ctx* c);  // slot-order planes, refreshed after an LHS update
    void ensure_gstif_tiles(cbs_hip_ctx* c);  // tile-order planes, the same
    void launch_constraints(cbs_hip_ctx* c);

    // k_sell.hip: the assembled operator of the PCG apply
    struct SellSource
    {
        int N = 0, F = 0, ntiles = 0, nseg = 1;
        const int32_t* tile_node_ptr = nullptr;  // element tiles (ntiles+1)
        const int32_t* tile_nodes = nullptr;
        const int32_t* seg_tile = nullptr;       // nseg+1 element-tile offsets
        const int32_t* seg_node = nullptr;       // nseg+1 node offsets
        const int4* conn = nullptr;              // element order, 0-based nodes
        const int32_t* inc_ptr = nullptr;        // node incidences (e*4 + a)
        const int32_t* inc = nullptr;
        const int32_t* elem_tslot = nullptr;     // tile slot of each element, -1 for solids
        const uint8_t* nflag = nullptr;
        int lds_bytes = 65536;                   // staging budget of one workgroup
    };
    struct SellHost
    {
        std::vector<int32_t> rows, row_base, row_cnt, halo_ptr{0}, halo, slice_ptr{0}, seg_tile{0}, gid;
        std::vector<int64_t> slice_off{0};
        std::vector<uint32_t> cols, cptr{0}, cidx;
        int64_t nnz = 0;
        int32_t lmax = 0, max_slices = 0;
    };
    bool build_sell_host(const SellSource& s, SellHost& out, std::string& err);
    void sell_configure(cbs_hip_ctx* c);
    void sell_refresh_values(cbs_hip_ctx* c);  // values from gstif_t
    // x_odd, iter0: x is read from x_odd in the segments whose iteration
    // count (iter0[seg * stride], a DevScalars field) is odd
    void sell_apply64(cbs_hip_ctx* c, const double* x, double* y, bool ap, const int* flag0,
                      const double* x_odd = nullptr, const int* iter0 = nullptr);
    // The FP32 applies use the difference form y_i = sum_j a_ij (x_j - x_i) +
    // rs_i x_i; `pdiagf` (rs) comes from sell_rowsum32 (ones, acc: N-double
    // scratch vectors; needs the solver-order pdiag).
    void sell_apply32(cbs_hip_ctx* c, const float* x, float* y, const float* pdiagf, const int* flag0,
                      const float* x_odd = nullptr, const int* iter0 = nullptr);
    // The PCG iteration's apply with the direction update fused into its
    // staging: iteration j >= 1 (the segment's DevScalars iter) forms p_j =
    // free ? z + beta p_{j-1} : 0 from p_{j-1} in the other buffer (p0 even
    // j, p1 odd j), writes it and applies the operator to it; iteration 0
    // applies p0 as it is.  z = r / pdiag (precond 1) or r in FP64, z = r
    // dinv rounded to FP32 in the mixed form.  force: form the direction at
    // iteration 0 too (for timing on scratch vectors).
    void sell_apply64_dir(cbs_hip_ctx* c, double* p0, double* p1, const double* r, int precond, double* y,
                          const int* flag0, int force = 0);
    void sell_apply32_dir(cbs_hip_ctx* c, float* p0, float* p1, const float* r, const float* dinv,
                          const float* pdiagf, float* y, const int* flag0, int force = 0);
    void sell_rowsum32(cbs_hip_ctx* c, double* ones, double* acc, float* rs);
    void sell_autotune64(cbs_hip_ctx* c, const double* x, double* y);
    void sell_autotune32(cbs_hip_ctx* c, const float* x, float* y, const float* pdiagf);
    void sell_set_config(cbs_hip_ctx* c, int form, int threads, int batch);
    // y = free ? A x : 0 on device vectors (FP64, or FP32 through f_p/f_ap/f_z)
    void sell_apply_test(cbs_hip_ctx* c, int precision, const double* x, double* y);
    // node order <-> solver order copies of an N-vector
    void sol_gather(cbs_hip_ctx* c, const double* node_order, double* solver_order);
    void sol_scatter(cbs_hip_ctx* c, const double* solver_order, double* node_order);
    // Multi-column products on multi-column vectors (mcol_at layout) in
    // solver order.  sell_spmm_groups builds the product's tiles (first call)
    // and picks the columns per workgroup (the largest whose staging fits
    // the LDS, at most 2^rshift; 0 if none) and allows its LDS; pass it to
    // the products.  ap: y = free ? A x : 0 with per-tile
    // column partials of (x, y) in part[(tile << rshift) + col]; otherwise y
    // = the off-diagonal product.  Nothing runs while *stop != 0; with a
    // device mask skip, column groups whose columns all have skip[col] != 0
    // are not computed.
    int sell_spmm_groups(cbs_hip_ctx* c, bool fp32, int rshift);
    int sell_spmm_tiles(const cbs_hip_ctx* c);  // workgroups per column group (the partials' tile count)
    void sell_spmm64(cbs_hip_ctx* c, int groups, const double* x, double* y, bool ap, double* part, int rshift,
                     const int* stop, const int* skip = nullptr);
    void sell_spmm32(cbs_hip_ctx* c, int groups, const float* x, float* y, const float* pdiagf, double* part,
                     int rshift, const int* stop, const int* skip = nullptr);
    void sell_gather_columns(cbs_hip_ctx* c, const double* node_order, double* solver_order, int rshift);
    void sell_scatter_columns(cbs_hip_ctx* c, const double* solver_order, double* node_order, int rshift);

    // k_adjoint.hip
    void destroy_adjoint(cbs_hip_ctx* c);

    // k_amg.hip: the multigrid-preconditioned solve
    int solve_pressure_amg(cbs_hip_ctx* c, cbs_cg_result* res);
    void destroy_amg(cbs_hip_ctx* c);
    double amg_setup_seconds(const cbs_hip_ctx* c);
    int amg_levels(const cbs_hip_ctx* c);
    int64_t amg_stored(const cbs_hip_ctx* c);
    void launch_step1_update_forced(cbs_hip_ctx* c);

    // launchers implemented in k_multi.hip
    int solve_pressure_multi(cbs_hip_ctx* c, int nrhs, const double* B, double* X, const cbs_cg_options* o,
                             cbs_cg_result* res);

    // apply timing helpers (k_pressure.hip)
    bool timer_begin(cbs_hip_ctx* c, int precision, int idx, int64_t seq);
    void timer_end(cbs_hip_ctx* c, bool begun);
    // after the stream is synchronised; drops pairs with idx >= keep_below
    void timer_collect(cbs_hip_ctx* c, int keep_below = 0x7fffffff);

    // Phases of ConjugateGradient::Result's timers (ConjugateGradient.hpp:55-62).
    enum CgPhase : int
    {
        PH_MATVEC = 0, PH_DOT = 1, PH_PRECOND = 2, PH_VECTOR = 3, PH_CONSTRAINT = 4
    };
    // phase_mark(c, -1, idx) opens a group of a sampled iteration (idx: its
    // position in the chunk) or, with idx < 0, a group timed every time;
    // phase_mark(c, ph) closes the kernels since the last mark as phase ph.
    void phase_mark(cbs_hip_ctx* c, int label, int idx = -1);
    void phase_reset(cbs_hip_ctx* c);
    void phase_collect(cbs_hip_ctx* c, int keep_below = 0x7fffffff);
    // After the solve's last launch: the Result timers of `res` (device time
    // per phase, sampled iterations scaled to res->iterations) and setup.
    void phase_fill(cbs_hip_ctx* c, cbs_cg_result* res, double setup_seconds);

    // The marker handed to the launch functions of one iteration: records
    // only for the iteration that carries the timing.
    struct PhaseMark
    {
        cbs_hip_ctx* c = nullptr;
        bool on = false;
        void operator()(int ph) const
        {
            if (on) phase_mark(c, ph);
        }
    };
    void destroy_graphs(cbs_hip_ctx* c);

    // Runs fn(i) for every i in [0, n) on up to 16 host threads, handing out
    // items one at a time; the first exception is rethrown after the join.
    template <typename F>
    void host_parallel_for(int64_t n, F&& fn)
    {
        const int64_t nth = std::min<int64_t>(n, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        if (nth <= 1)
        {
            for (int64_t i = 0; i < n; ++i) fn(i);
            return;
        }
        std::atomic<int64_t> next{0};
        std::exception_ptr first;
        std::mutex mu;
        std::vector<std::thread> pool;
        for (int64_t t = 0; t < nth; ++t)
            pool.emplace_back([&] {
                try
                {
                    for (int64_t i = next++; i < n; i = next++) fn(i);
                }
                catch (...)
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (!first) first = std::current_exception();
                    next = n;
                }
            });
        for (auto& th : pool) th.join();
        if (first) std::rethrow_exception(first);
    }

    // k_setup.hip: Coloring::build on the device; colour_elements takes host
    // arrays and allocates its workspace per call, colour_elements_noalloc
    // takes device arrays and works in a fixed device array
    int colour_elements_noalloc(int32_t nelem, int32_t npoin, const int32_t* intma_d, int32_t* ncolor,
                                int32_t* color_ptr_d, int32_t* color_elem_d, int32_t* levels, std::string& msg);
    int64_t colour_noalloc_scratch_bytes();
    int colour_elements(int32_t nelem, int32_t npoin, const int32_t* intma, int32_t* ncolor, int32_t* color_ptr,
                        int32_t* color_elem, int32_t* levels, std::string& msg);

    // ctx.hip
    // The CSR pattern of the pressure operator (built on first use).
    int check_device_error(cbs_hip_ctx* c);
    int set_error(cbs_hip_ctx* c, int status, const std::string& msg);
    void hip_check(hipError_t e, const char* what);
}
