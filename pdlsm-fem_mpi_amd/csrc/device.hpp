// Device context of the PDLSM-FEM hot path on one MI355X (gfx950).
//
// Layout in HBM
//   * stiffness: node-block CSR ("BSR") over the node graph of the reference's
//     pattern (every neighbour node of node k, prescribed DOFs included), with
//     d x d blocks stored contiguously: value (i,j) of block b lives at
//     K[b*d*d + i*d + j], so one lane loads a whole block with 2-5 wide loads
//     (dwordx4) and the lanes of a row segment read consecutive blocks.
//   * vectors: full node-major DOF vectors of length nn*d; prescribed DOFs
//     are masked where a free-DOF quantity is meant.
//   * families: CSR of member node ids (self first) with the block index of
//     every member in the family node's row (the scatter map: one store per
//     block, no atomics).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <functional>
#include <cstdint>
#include <string>
#include <vector>

#include "model.hpp"

namespace pdx {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// explicit-step kernel forms searched by the tuner: lanes per short row, non-temporal K loads,
// XCD-contiguous block order, row-local value order of K, column pattern classes, mirrored
// lower blocks read from their transposed partners, and the sweep direction (alt: odd steps
// visit the blocks in reverse order, so a step starts on the rows the previous step read last,
// which are still in the memory-side cache when the step's data exceeds it; the same kernel as
// form `base`, same bits)
struct ExplicitForm {
    int lanes;
    bool nt, xm, soa, rel, sym, alt;
    int base;
};
constexpr ExplicitForm kExplicitForms[PDX_EXPLICIT_FORMS] = {
    {8, false, false, false, false, false, false, 0},  {4, false, false, false, false, false, false, 1},
    {16, false, false, false, false, false, false, 2}, {8, true, false, false, false, false, false, 3},
    {4, true, false, false, false, false, false, 4},   {16, true, false, false, false, false, false, 5},
    {32, false, false, false, false, false, false, 6}, {32, true, false, false, false, false, false, 7},
    {16, false, true, false, false, false, false, 8},  {8, false, true, false, false, false, false, 9},
    {8, false, true, true, false, false, false, 10},   {16, false, true, true, false, false, false, 11},
    {4, false, true, true, false, false, false, 12},   {32, false, true, true, false, false, false, 13},
    {8, false, true, false, true, false, false, 14},   {16, false, true, false, true, false, false, 15},
    {8, false, true, true, true, false, false, 16},    {16, false, true, true, true, false, false, 17},
    {8, false, true, false, true, true, false, 18},    {16, false, true, false, true, true, false, 19},
    {16, false, true, false, true, true, true, 19},    {8, false, true, false, true, true, true, 18},
    {16, false, true, true, true, false, true, 17},    {16, false, true, false, true, false, true, 15}};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the K assembly (the family kernel beside the element kernel) and its
    // fork / join events; created on the first assembly
    hipStream_t side = nullptr;
    int *asm_status = nullptr;  // status word of the assembly kernels
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    const Model *host = nullptr;  // not owned; copied fields below
    int dim = 3, dd = 9, nn = 0, ne = 0, neq = 0, nfam = 0, max_family = 0, npdbe = 0;
    long long nblk = 0, nnz = 0;
    double lambda = 0, mu = 0, nlf = 0.33, rho = 0, dt = 1, beta = 0.25, gamma = 0.5;

    // geometry / model
    double *x = nullptr, *nv = nullptr, *ev = nullptr;
    double *Dm = nullptr;  // [36]
    int *econ = nullptr, *nen = nullptr;
    int *eq = nullptr, *presc_set = nullptr;  // [nn*3]
    int *fam_node = nullptr, *fam_off = nullptr, *fam_mem = nullptr, *bond = nullptr, *fam_blk = nullptr;
    double *fam_v = nullptr, *fam_delta = nullptr;
    double *fam_c = nullptr;  // [members*d] strain weights of the bonds (written by the family kernel)
    int *fp_off = nullptr, *fp_list = nullptr, *pdbe = nullptr;
    int max_pdbe_per_family = 0;  // most PD-boundary entities incident to one family
    int nfe = 0;
    int *fe_elem = nullptr, *fe_blk = nullptr, *nfe_off = nullptr, *nfe_item = nullptr;
    // element block staging: fe_slot[t * NP + pair] = (staging slot << 1 | transposed) of the
    // upper node-pair block `pair` of FE element t (-1 for pairs the element does not have); the
    // slots of one K node pair are consecutive, in element order (pair_off)
    int *fe_slot = nullptr, *pair_off = nullptr;
    long long n_stage = 0;
    // node pairs of the gather: upper K blocks (r, c), r <= c, with FE terms, and their lower
    // partners (c, r) (-1 on the diagonal); the first n_fe_pair_fe pairs have no PD part in
    // either block
    int n_fe_pair = 0, n_fe_pair_fe = 0;
    int *fe_pair_up = nullptr, *fe_pair_lo = nullptr;
    int *fe_pair_row = nullptr;  // row node r of each pair with a PD part (pairs n_fe_pair_fe ...)
    unsigned char *pd_bits = nullptr;        // per K block: 1 if a family kernel stores it
    bool k_needs_fill = true;                // some K block is written by neither assembly kernel
    int n_fe_a = 0, n_fe_b = 0;  // FE-list positions of hex8/quad4 (a) and tet4/tri3 (b) elements
    int *fe_list_a = nullptr, *fe_list_b = nullptr;
    double *Ke_stage = nullptr;  // element stiffness staging [nfe][ND*ND]
    int *nall_off = nullptr, *nall_ele = nullptr;
    int *adj_off = nullptr, *adj = nullptr;
    int64_t *ia = nullptr;
    int *row_node = nullptr;  // [neq] node*3+dof of each free row
    // rows of the node graph binned by length for the SpMV (short: one group of 8 lanes, long: a wave)
    int n_short = 0, n_long = 0;
    unsigned spmv_calls = 0;  // products of the 3-DOF Krylov SpMV so far (their sweep alternates)
    int *rows_short = nullptr, *rows_long = nullptr;

    // natural BCs
    int n_nbc = 0, n_nbc_ele = 0;
    int *nbc_con = nullptr, *nbc_set = nullptr;
    double *nbc_fe = nullptr;     // unit-traction element forces [n_nbc_ele * 12]
    int *nbcn_off = nullptr, *nbcn_item = nullptr;  // node -> (element*4+local)
    double *f_unit = nullptr;     // [n_nbc * nn*d] unit-traction nodal forces
    std::vector<double> nbc_val, nbc_vel;
    std::vector<int> nbc_vary;
    std::vector<double> ebc_val, ebc_vel;
    std::vector<int> ebc_vary;
    double *d_ebc_vel = nullptr;   // [n_ebc] per-step increment dt*vel (0 if constant)
    double *d_nbc_vel = nullptr;   // [n_nbc]

    // matrices
    double *K = nullptr;   // [dd*nblk]
    float *K32 = nullptr;  // [dd*nblk]
    // K in row-local value order for the explicit step: value e of block b of the row whose blocks
    // are [b0, b0 + nb) at b0*dd + e*nb + (b - b0), so that lanes on consecutive blocks read
    // consecutive words; built on first use for the K of generation ks_gen
    // column pattern classes of the node rows: row r's block columns are r + cls_rel[cls_off[cls[r]] + k]
    int *cls = nullptr, *cls_off = nullptr, *cls_rel = nullptr;
    int n_cls = 0;
    // mirrored storage of K for the explicit step: the "own" blocks (diagonal, upper, and lower
    // blocks with a PD part in them or in their partner) in row order, and per K block its own
    // index, or ~(own index of its partner) for a lower block read as the partner's transpose
    int *ksrc = nullptr, *kown_slot = nullptr;
    long long n_own = 0;
    double *Kown64 = nullptr;
    float *Kown32 = nullptr;
    long long kown_gen[2] = {-1, -1};
    long long kown_exact_gen = -1;  // K generation whose mirrored blocks were checked exact
    bool kown_exact = false;
    double *K64s = nullptr;
    float *K32s = nullptr;
    long long ks_gen[2] = {-1, -1};
    double *Mk = nullptr;  // consistent mass on K blocks (Newmark)
    bool K_valid = false, K32_valid = false, Mk_valid = false;
    long long K_gen = 0;  // incremented by every assembly of K
    // bond changes since K was assembled: the families whose bond states changed, applied to K
    // by recomputing those families' rows (K_full: K holds a whole assembly; incremental_bonds:
    // off, every change assembles K again from scratch)
    std::vector<int> dirty_fams;
    bool K_full = false, incremental_bonds = true;
    int *d_flist = nullptr;
    unsigned char *d_rowflag = nullptr;
    double *mlump = nullptr;  // [nn*d] lumped mass (1 for prescribed)
    bool mlump_valid = false;

    // explicit state: three rotating full DOF vectors, two NBC value buffers
    double *U[3] = {nullptr, nullptr, nullptr};
    double *nbcv = nullptr;  // [2 * n_nbc]
    int step_phase = 0;      // buffer rotation counter
    long long steps_done = -1;  // -1: start-up step not yet taken
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // per precision
    int tune_explicit[2] = {-1, -1};               // chosen step-kernel variant per precision
    float tune_ms[2][PDX_EXPLICIT_FORMS] = {{0}};  // measured ms per variant (-1: not run)
    // consistent-mass explicit steps (LUMPMASS OFF): F = F_ext - K u_n and the acceleration of
    // M a = F, full DOF vectors; the acceleration warm-starts the next step's solve
    double *nlm_F = nullptr, *nlm_a = nullptr;
    long long nlm_iters = 0;  // CG iterations of the mass solves so far
    double *dofs_tmp = nullptr;  // [nn*d] DOF values of the last stored step (stress input)

    // scratch for Krylov
    std::vector<DevBuf> scratch;
    double *red = nullptr;      // partial sums [RED_SLOTS * RED_PARTS]
    double *red_host = nullptr; // pinned mirror
    double last_total_ms = 0, last_step_ms = 0;
    void *two_level = nullptr;       // coarse space of the last two-level solve (krylov.hip)
    void *ilu0 = nullptr;            // block ILU(0) factor of the last ILU(0) solve (ilu0.hip)
    void *krylov_ws[2] = {nullptr, nullptr};  // Krylov workspaces (FP64, FP32) and their graphs (krylov.hip)
    double *solve_y = nullptr;   // FP64 residual of the refinement / restart loop [nn*d]
    float *solve_b32 = nullptr;  // FP32 right-hand side of the refined inner solves [nn*d]
    double *learned_Minv = nullptr;  // [nn*dd] node-block smoother predicted by the graph network
    long long learned_gen = -1;      // K_gen the smoother was predicted for
    const double *learned_for = nullptr;  // the operator it was predicted for (K, or a Newmark K_eff)
    // operator the network's inputs, training probes and pdx_gnn_apply use: 0 K, 1 the Newmark
    // operator M + beta dt^2 K (formed in gnn_keff for the current K); the network a learned
    // Newmark solve predicts its smoother with
    int gnn_op = 0;
    double *gnn_keff = nullptr;
    long long gnn_keff_gen = -1;
    const void *gnn_net = nullptr;
    void *gnn_work = nullptr;        // graph-network inputs and activations (learned.hip)
    double *stress_stage = nullptr;  // [nfe*8*6 + nfam*6] element corner stresses, family stresses
    int *stress_famid = nullptr;     // [nn] family id of every node (-1: none)
    int *stress_status = nullptr;    // singular-Jacobian flag of the stress recovery

    // helpers
    std::vector<void *> owned;
};

// process-wide serial numbers of rebuilt device objects (a rebuild may reuse freed addresses)
inline long long next_build_serial() {
    static std::atomic<long long> n{0};
    return ++n;
}

constexpr int RED_PARTS = 512;  // blocks of every reducing kernel
constexpr int RED_SLOTS = 16;

int ctx_build(const Model &m, int device, Ctx *&out, std::string &err);
void ctx_free(Ctx *c);

// assembly (assemble.hip)
int dev_assemble_K(Ctx &c, std::string &err);
int dev_assemble_K_launch(Ctx &c, std::string &err);  // enqueue only
int dev_assemble_K_finish(Ctx &c, std::string &err);  // wait, check the status word
int dev_lumped_mass(Ctx &c, std::string &err);
int dev_nbc_unit(Ctx &c, std::string &err);
int dev_mass_on_K(Ctx &c, std::string &err);
int dev_extract_csr(Ctx &c, const double *Kblk, double *Ku_dev);
int dev_spmv(Ctx &c, const double *x, double *y, int prec, bool mask_free);
int dev_fill(Ctx &c, double *p, double v, size_t n);

int dev_spmv_K(Ctx &c, const double *K64, const double *x, double *y);
// Own-block copies (mirrored storage, ksrc order) of one node-block operator for the Krylov
// SpMVs; a null member means that precision reads the full operator
struct Mirrored {
    const double *K64 = nullptr;
    const float *K32 = nullptr;
};
// exact: every mirrored lower block of A (dd*nblk values, K's block order) equals its partner's
// transpose bit for bit (checked on the device; synchronises the stream)
int check_mirrored(Ctx &c, const double *A, bool &exact, std::string &err);
// own64 / own32 (either may be null; dd*n_own values) <- the own blocks of A
int gather_own(Ctx &c, const double *A, double *own64, float *own32, std::string &err);
// mirrored copies of the context's K for a solve in precision prec (empty when not exact)
int kown_for_solve(Ctx &c, int prec, Mirrored &m, std::string &err);
int dev_solve_mat(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                  double *x_dev, pdx_solve_stats &st, std::string &err, const Mirrored *mir = nullptr);
void free_two_level(Ctx &c);
// block ILU(0) preconditioner (ilu0.hip): factor of K64 cached per generation, applications on c.stream
int ilu0_prepare(Ctx &c, const double *K64, bool f32, void *&out, std::string &err);
void ilu0_apply(Ctx &c, const void *h, const double *r, double *z);
void ilu0_apply(Ctx &c, const void *h, const float *r, float *z);
int ilu0_colors(const void *h);
long long ilu0_serial(const void *h);
double ilu0_setup_ms(const void *h);
void free_ilu0(Ctx &c);
void free_krylov_ws(Ctx &c);
void free_gnn_work(Ctx &c);
// smoother of the context's attached network for operator A (learned.hip)
int dev_learned_smoother(Ctx &c, const double *A, std::string &err);
void forget_learned_operator(Ctx &c, const double *A);
// K_eff = beta dt^2 K + M (newmark.hip)
int dev_form_keff(Ctx &c, double *Keff, std::string &err);
// nodal stresses of u_dev; check: read the singular-Jacobian flag (synchronises the stream)
int dev_node_stresses(Ctx &c, const double *u_dev, double *sigma_dev, std::string &err, bool check = true);
// out[t] = free_src[t] at free DOFs, presc_src[t] at prescribed ones (full DOF vectors, solve.hip)
int dev_merge_dofvals(Ctx &c, const double *free_src, const double *presc_src, double *out);
int stress_status_check(Ctx &c, std::string &err);
int dev_bond_sigma1(Ctx &c, const int *bond_dev, const double *sigma_dev, double *s1_dev, std::string &err);
double two_level_setup_ms(Ctx &c);
int set_error(int rc, const std::string &msg);

}  // namespace pdx

// handle types behind the opaque C-ABI pointers of include/pdx.h
struct pdx_model {
    pdx::Model m;
};
struct pdx_ctx {
    pdx::Ctx *c = nullptr;
    const pdx_model *model = nullptr;
    std::vector<double> dofval;  // DOF values used by the assembly modes, [nn*dim]
    std::vector<int> bond;       // bond states of this context (the model's, then broken by failure)
    // per stored step of the dynamic drivers with FC 2: bonds kept, bonds broken, largest s1
    std::vector<double> step_log;
};

#define PDX_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return PDX_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)
