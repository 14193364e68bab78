/*
 * pdx.h -- C-ABI of the MI355X PDLSM-FEM hot-path library (libpdx.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (PDX_OK = 0, negative on error) instead of calling exit() as the reference
 * does; pdx_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Reference seams (QibangLiu/PDLSM-FEM_mpi) each entry point replaces:
 *   pdx_model_create / pdx_model_read
 *       datModel::readdata            SourceFiles/Solver/datModel.cpp:58-326
 *       fioFiles::excuteCMD           SourceFiles/Solver/fioFile.cpp:69-462
 *       pdsolve::setDatModel          SourceFiles/Solver/pdsolve.cpp:114-151
 *       pdsolve::setCSRIndexes_gloStiffMat  pdsolve.cpp:959-1121
 *   pdx_assemble
 *       pdsolve::assembleInterWorkPD_CSRformat  pdsolve.cpp:2524-2730 (HeaderFiles/Solver/pdsolve.h:100)
 *       pdsolve::assemblePDBEwork_CSRformat     pdsolve.cpp:3674-4273 (pdsolve.h:107)
 *       pdsolve::assembleSEDbyFEM_CSRformat     pdsolve.cpp:4882-5007 (pdsolve.h:122)
 *       pdsolve::calExternalForce_CSRformat     pdsolve.cpp:5009-5071 (pdsolve.h:123)
 *   pdx_lumped_mass
 *       pdsolve::assembleLumpedMass             pdsolve.cpp:4618-4669
 *   pdx_explicit_run / pdx_explicit_steps
 *       pdsolve::calinternalForce_CSRformat     pdsolve.cpp:5819-5840 (pdsolve.h:128)
 *       pdsolve::timeIntegration_LM             pdsolve.cpp:5113-5132 (pdsolve.h:145)
 *       driver loop of pdfemDynamicSolver_CSRformat  pdsolve.cpp:5842-6090
 *   pdx_static_solve
 *       pdsolve::pdfemStaticSolver_CSRformat    pdsolve.cpp:5557-5654
 *   pdx_newmark_run
 *       pdsolve::pdfemDynamicNewmarkSolver_CSRformat pdsolve.cpp:6092-6290,
 *       pdfemAssembleKNFR_CSRformat :6327-6397, updateDispVelo :6399-6422
 *   pdx_solve_csr
 *       calMatrixOperations::cluster_PARDISO_64Solver
 *       SourceFiles/Solver/calMatrixOperations.cpp:378-498
 *       (HeaderFiles/Solver/calMatrixOperations.h:47-48)
 */
#ifndef PDX_H
#define PDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define PDX_OK 0
#define PDX_ERR_ARG (-1)
#define PDX_ERR_SINGULAR (-2)
#define PDX_ERR_PATTERN (-3)
#define PDX_ERR_MESH (-4)
#define PDX_ERR_IO (-5)
#define PDX_ERR_HIP (-6)
#define PDX_ERR_NOCONV (-7)
#define PDX_ERR_LIMIT (-8)

/* assembly modes: the reference's ci_solvFlag / cb_Newmark (datModel.h:122, :132) */
#define PDX_MODE_EXPLICIT 0 /* solvFlag 0, Newmark off: F = F_ext - K u          */
#define PDX_MODE_STATIC 1   /* solvFlag 1: K_ff, F = F_ext - K_fp u_p            */
#define PDX_MODE_QUASI 2    /* solvFlag 2: K_ff, F = F_ext - K u (all columns)   */
#define PDX_MODE_NEWMARK 3  /* solvFlag 0, Newmark on: K_ff, F = F_ext - K U_N   */

/* precision of the explicit internal force / Krylov inner iterations */
#define PDX_FP64 0
#define PDX_FP32 1 /* FP32 K values, FP32 accumulation, FP64 state */

/* Krylov methods and preconditioners for pdx_static_solve / pdx_solve_csr */
#define PDX_KRYLOV_BICGSTAB 0
#define PDX_KRYLOV_CG 1
#define PDX_PREC_NONE 0
#define PDX_PREC_JACOBI 1
#define PDX_PREC_BLOCK_JACOBI 2 /* node d x d blocks */
#define PDX_PREC_LEARNED 3      /* two-level, node-block smoother predicted by the graph network */
#define PDX_PREC_TWO_LEVEL 4    /* two-level hybrid: rigid-body aggregate coarse space + node-block Jacobi */
#define PDX_PREC_ILU0 5         /* node-block ILU(0) in multicolour order (BiCGStab only) */

typedef struct pdx_model pdx_model; /* host-side data model (families, pattern) */
typedef struct pdx_ctx pdx_ctx;     /* device-resident state on one GPU         */

/* Raw mesh + solver settings: the content of a .lis mesh file after
 * datModel::readdata plus the SETSOLVING/EBC/VEBC/NBC/VNBC commands.
 * Arrays are caller-owned and copied by pdx_model_create. */
typedef struct pdx_mesh_desc {
    int dim;      /* 2 or 3 */
    int pro_type; /* 2D: 1 plane stress, 2 plane strain */
    double E, nu, rho;
    int n_node;
    const double *xyz; /* [n_node*3] */
    int n_ele;
    const int *ele_nodes; /* [n_ele*8] 1-based node ids */
    const int *ele_nen;   /* [n_ele] 3/4 (2D), 4/8 (3D) */
    const int *ele_algo;  /* [n_ele] 1 = PD element, 2 = FE element */
    int n_ebc;
    const int *ebc_off;   /* [n_ebc+1] */
    const int *ebc_nid;   /* [ebc_off[n_ebc]] 1-based */
    const int *ebc_dof;   /* [n_ebc] 0 UX, 1 UY, 2 UZ */
    const double *ebc_val, *ebc_vel;
    const int *ebc_vary;
    int n_nbc;
    const int *nbc_off;   /* [n_nbc+1] elements per set */
    const int *nbc_con;   /* [nbc_off[n_nbc]*(dim==3?4:2)] 1-based */
    const double *nbc_val, *nbc_vel;
    const int *nbc_vary;
    int n_crack;
    const double *crack;  /* [n_crack*9] three points (2D: two) */
    double horizon_factor; /* m of SETSOLVING */
    double nlf;            /* non-local factor a */
    double me;             /* enrichment radius factor */
    double dt, beta, gamma;
    double sigma_ult, kic; /* material strength and toughness (mesh file) */
    int fc, topk;          /* failure criterion and bonds per step (FC / TOPK commands) */
    /* run settings of the command file; zero is the reference's default (datModel.cpp:17-56) */
    int lumped_mass;       /* LUMPMASS ON (1) / OFF (0, default): explicit driver's mass */
    int newmark;           /* NEWMARK ON (1) / OFF (0, default) */
    int save_freq;         /* SETSOLVING token 3; <= 0 means 1 */
    int vtk_ascii;         /* VTKFORMAT ASCII (1) / BINARY (0, default) */
} pdx_mesh_desc;

/* Settings of a model read from a command file (fioFiles::excuteCMD, fioFile.cpp:69-462) or
 * given in a pdx_mesh_desc, as pdx_model_settings reports them. */
typedef struct pdx_settings {
    int solver_flag;   /* SOLVER: 0 DYNAMIC, 1 STATIC, 2 QUASI-STATIC, -1 none (pdx_model_create) */
    int newmark;       /* NEWMARK ON|OFF: with solver_flag 0, Newmark (1) or central difference (0) */
    int lumped_mass;   /* LUMPMASS ON|OFF */
    int num_steps;     /* SETSOLVING token 2 */
    int save_freq;     /* SETSOLVING token 3: result files every save_freq steps and at the last */
    int vtk_binary;    /* VTKFORMAT: 1 BINARY (default), 0 ASCII */
    int fc, topk;      /* FC flag [TopK], TOPK */
    int n_rf;          /* RF: number of EBC set ids (the first 16 in rf_ids) */
    int rf_ids[16];
    double dt, beta, gamma, horizon_factor, nlf;
    double sigma_ult, kic;  /* material strength and toughness of the mesh file */
} pdx_settings;

typedef struct pdx_solve_opts {
    int krylov;        /* PDX_KRYLOV_* */
    int precond;       /* PDX_PREC_* */
    int precision;     /* PDX_FP64: plain FP64 Krylov; PDX_FP32: FP32 inner + FP64 refinement */
    double tol;        /* true relative residual target, e.g. 1e-8 */
    int max_iter;      /* total inner iterations */
    double inner_tol;  /* FP32 inner relative tolerance (refinement), e.g. 1e-4 */
    int warm_start;    /* use x as the initial guess; Newmark: 0/1 start from a_n, 2 from 2 a_n - a_{n-1} */
    int coarse_agg;    /* two-level: aggregate edge in node spacings (0: 10 in 2D, 6 in 3D) */
    int fallback;      /* nonzero: when the solve stops short of tol (PDX_ERR_NOCONV), a preconditioner is
                          singular, or PDX_PREC_LEARNED has no smoother for this operator, retry with
                          BiCGStab down the chain learned / two-level -> ILU(0) -> node-block Jacobi
                          (any other preconditioner -> node-block Jacobi) -> FP64 node-block Jacobi,
                          from the iterate the failed attempt left when it beat zero */
    int no_graph;      /* nonzero: launch every Krylov iteration's kernels one by one instead of
                          replaying captured hipGraphs of 10 iterations (same results, bit for bit) */
    int mirrored;      /* nonzero (2D / 3D models): the SpMVs read the operator from mirrored storage,
                          each lower block that no family stores in either orientation read as its
                          upper partner's transpose, once a device check has found every such block
                          equal to that transpose bit for bit (else full storage; same results, bit
                          for bit, either way) */
} pdx_solve_opts;

typedef struct pdx_solve_stats {
    int iterations;    /* inner Krylov iterations */
    int refinements;   /* FP64 refinement sweeps */
    double true_relres; /* ||b - A x|| / ||b|| in FP64 */
    double seconds;
    int converged;
    int precond_used;   /* PDX_PREC_* of the attempt that ended the solve */
    int precision_used; /* PDX_FP64 / PDX_FP32 of that attempt */
    int attempts;       /* 1 + fallback retries */
    int mirrored_used;  /* 1 when the SpMVs read mirrored storage (pdx_solve_opts.mirrored) */
} pdx_solve_stats;

const char *pdx_last_error(void);
const char *pdx_version(void);

/* ---- host data model -------------------------------------------------- */
/* Where later pdx_model_create / pdx_model_read calls run the family search (initialPDN_Fami,
 * pdsolve.cpp:1442-1568) and the node graph of the stiffness pattern (setCSRIndexes_gloStiffMat,
 * pdsolve.cpp:959-1121): -1 (the default) on the host, threaded; device >= 0 on that GPU.  Both
 * give the same families and pattern.  Process-wide; PDX_ERR_ARG for a device that does not
 * exist. */
int pdx_set_setup_device(int device);
int pdx_model_create(const pdx_mesh_desc *desc, pdx_model **out);
int pdx_model_read(const char *cmd_file, pdx_model **out, int *solver_flag, int *num_steps);
void pdx_model_destroy(pdx_model *m);
/* the model's run settings: what fioFiles::excuteCMD (fioFile.cpp:69-462) leaves in datModel
 * (solver flag, NEWMARK, LUMPMASS, SETSOLVING, FC / TOPK, RF, VTKFORMAT) */
int pdx_model_settings(const pdx_model *m, pdx_settings *out);
/* Pattern of the explicit driver's consistent mass, pdsolve::setCSRIndexes_gloMassMat
 * (pdsolve.cpp:5674-5817): zero-based int64 CSR over the free DOFs, sorted columns; rows of
 * pure PD nodes hold their diagonal only.  ia [neq+1] and ja [*nnz] may be NULL (size query). */
int pdx_model_mass_csr(const pdx_model *m, long long *ia, long long *ja, long long *nnz);
/* Result file of the whole mesh in the reference's legacy VTK layout, byte for byte:
 * writeReslutsTOTAL_vtk_Binary / _ASCII (fioFile.cpp:562-769), chosen as its VTKFORMAT
 * command does (binary: big-endian).  u [n_node*3] DOF values (UZ 0 in 2D), sigma [n_node*6]
 * nodal stresses (xx, yy, zz, xy, yz, zx, as pdx_node_stresses gives them), phi [n_node]
 * nodal damage; a NULL array writes zeros.  title NULL: the mesh file's title. */
int pdx_write_vtk(const pdx_model *m, const char *path, const char *title, const double *u, const double *sigma,
                  const double *phi, int binary);
/* The ASCII files of one part of the mesh that the static driver writes beside the whole-mesh
 * file (fioFiles::writeResults with its default flag 3, fioFile.cpp:464-482): part 0
 * writeResultsPD_vtk (PD nodes and PD elements, with phi), part 1 writeResultsFEM_vtk (nodes that
 * are not pure PD, FE elements), fioFile.cpp:771-1001.  Arrays as for pdx_write_vtk. */
int pdx_write_vtk_part(const pdx_model *m, const char *path, const char *title, const double *u, const double *sigma,
                       const double *phi, int part);
/* The node tables an ASCII whole-mesh result comes with (fioFile.cpp:486-538): kind 0 writeUofNode
 * (".disp": id, x, y, z, Ux, Uy, Uz), kind 1 writeSigofNode (".sig": id, x, y, z, six stresses). */
int pdx_write_node_table(const pdx_model *m, const char *path, const double *u, const double *sigma, int kind);
/* out[0..15]: n_node, n_ele, dim, neq, nnz, nfam, n_members, npdbe, npde, nfe,
 *             n_adj (node-graph entries), n_presc, nb0, nb1, nb2, max_family */
int pdx_model_sizes(const pdx_model *m, long long *out);
int pdx_model_csr(const pdx_model *m, long long *ia, long long *ja);
/* node graph of the pattern (setCSRIndexes_gloStiffMat pdsolve.cpp:992-1051 before
 * the DOF expansion): sorted neighbour nodes per node, prescribed DOFs included;
 * adj_off [n_node+1], adj [sizes[10]]; the order of the node blocks on the device */
int pdx_model_adjacency(const pdx_model *m, int *adj_off, int *adj);
/* families: node (0-based), offsets, members (1-based, self first), bond state, cd_v, horizon */
int pdx_model_families(const pdx_model *m, int *fam_node, int *fam_off, int *fam_mem, int *bond,
                       double *cd_v, double *delta);
/* eq [n_node*3], node type [n_node], node volume [n_node], element volume [n_ele],
 * boundary entities [npdbe*4] (1-based), element algorithm after crack update [n_ele] */
int pdx_model_nodes(const pdx_model *m, int *eq, int *ntype, double *nv, double *ev, int *pdbe, int *algo);
/* initial free-DOF displacement (from DOF values) and prescribed values, [n_node*dim] */
int pdx_model_values(const pdx_model *m, double *u_all);

/* ---- device context --------------------------------------------------- */
int pdx_ctx_create(const pdx_model *m, int device, pdx_ctx **out);
void pdx_ctx_destroy(pdx_ctx *c);
/* the HIP stream every kernel of this context runs on (hipStream_t) */
void *pdx_ctx_stream(pdx_ctx *c);

/* Assemble K and F on the GPU in the reference's CSR layout.
 * Ku: [nnz] K_ff values (NULL in PDX_MODE_EXPLICIT); F: [neq].
 * UN: [n_node*dim] predictor displacement for PDX_MODE_NEWMARK, else NULL.
 * with_ext: add the natural-BC force at the current NBC values. */
int pdx_assemble(pdx_ctx *c, int mode, const double *UN, double *Ku, double *F, int with_ext);
int pdx_lumped_mass(pdx_ctx *c, double *m);
/* Re-assemble K in device memory only (no host copies); ms: kernel time of the
 * family + element kernels measured with HIP events on the context stream. */
int pdx_assemble_timed(pdx_ctx *c, double *ms);
/* consistent FE mass + PD lumped mass on the K pattern, [nnz] */
int pdx_mass_on_K(pdx_ctx *c, double *Mv);
/* the same mass on its own pattern (pdx_model_mass_csr), [nnz of that pattern]:
 * assembleElemassMatFEM_CSRformat (pdsolve.cpp:5154-5210) + assembleMassMatPD_CSRformat
 * (:4568-4595) as the explicit driver builds it (:5923-5935) */
int pdx_mass_csr(pdx_ctx *c, double *Mv);

/* DOF values (free and prescribed, [n_node*dim]) used by PDX_MODE_EXPLICIT and
 * PDX_MODE_QUASI assembly and by pdx_static_solve for the prescribed part;
 * the reference keeps them in pdDof::cd_u (pdDof.h:18). */
int pdx_set_dof_values(pdx_ctx *c, const double *u_all);

/* Explicit central difference, pdfemDynamicSolver_CSRformat (pdsolve.cpp:5842-6090): the
 * start-up step u1 = u0 + dt^2/2 a0 on the first call, then nsteps loop steps, each advancing
 * the BCs before the force.  The model's LUMPMASS selects the mass: ON, a = F / m
 * (timeIntegration_LM :5113-5132); OFF (the reference's default), M a = F with the consistent
 * mass of pdx_mass_csr, solved on the device by PCG to a 1e-13 relative residual
 * (timeIntegration_NLM :5134-5152; the reference factorises M with PARDISO).  With FC 2 in 2D
 * every stored step ends with the nodal stresses and the bond failure of failureProcess
 * (:6062-6064, pdx_step_log); FC values other than 0 and 2 return PDX_ERR_ARG.
 * u_out/u_prev: [neq] host (may be NULL). */
int pdx_explicit_run(pdx_ctx *c, int nsteps, int precision, double *u_out, double *u_prev);
/* Same loop on device-resident state only, no host copies (for timing); lumped mass without
 * failure replays captured 6-step graphs when use_graph is set. */
int pdx_explicit_steps(pdx_ctx *c, int nsteps, int precision, int use_graph);
/* Per stored step of the dynamic drivers with FC 2 since the context was created (the
 * explicit start-up step included): bonds kept, bonds broken, largest principal bond stress;
 * info [*n * 3] (or NULL for the count). */
int pdx_step_log(const pdx_ctx *c, double *info, int *n);
/* Largest eigenvalue of M^-1 K (power iteration) -> critical dt = 2/sqrt(lmax) */
int pdx_max_eig(pdx_ctx *c, int iters, double *lmax);
int pdx_set_dt(pdx_ctx *c, double dt);

int pdx_static_solve(pdx_ctx *c, const pdx_solve_opts *o, double *u, pdx_solve_stats *st);

/* The block ILU(0) factor behind PDX_PREC_ILU0, for the context's assembled K (rows and
 * columns of prescribed DOFs removed, 1 on their diagonal).  No counterpart in the reference,
 * whose solves are MKL PARDISO (calMatrixOperations.cpp:378-498); IKJ ILU(0) over the d x d node
 * blocks, nodes eliminated by greedy colour (node order within a colour).  Each output may be
 * NULL: colors [n_node], F [blocks*d*d] in pdx_model_adjacency's block order (strict lower
 * blocks hold L, the rest U), Dinv [n_node*d*d] the inverses of U's diagonal blocks, and with r
 * [n_node*d] z = (L U)^-1 r. */
int pdx_ilu0(pdx_ctx *c, const double *r, double *z, int *colors, double *F, double *Dinv, int *ncolors);

/* Quasi-static load steps with bond failure: pdfemQuasiStaticSolver_CSRformat
 * (pdsolve.cpp:6464-6580).  Per step: K_ff du = F_ext - K u at the current DOF values (the
 * quasi-static assembly, pdsolve.cpp:6434-6462), u += du, the nodal stresses, failureProcess
 * (:6597-6641; 2D only, as in the reference): with fc == 2 the maximum-principal-stress
 * criterion breaks bonds (failureCriterion_maxPriSig and Finally_TopK_andUpdate_bondStaus,
 * :7184-7289, :6981-7182; K is assembled again from the new bond states), then, if no bond
 * broke, the natural and essential BCs advance (setVaryNaturalBC / setVaryEssentialBC with the
 * ratio 1.001 sigma_ult / max s1, :5263-5329).  fc, topk < 0 and sigma_ult <= 0 take the
 * model's FC / TOPK commands and material strength.  u_all [n_node*dim] gets the final DOF
 * values; step_info [nsteps*4] (or NULL) gets per step: bonds selected, bonds broken, ratio,
 * largest principal bond stress.  The context keeps its bond states (pdx_ctx_bonds) and DOF
 * values, so calls continue one another. */
int pdx_quasi_static_run(pdx_ctx *c, int nsteps, const pdx_solve_opts *o, int fc, int topk, double sigma_ult,
                         double *u_all, double *step_info, pdx_solve_stats *st);
/* the context's bond states, laid out like pdx_model_families' bond array */
int pdx_ctx_bonds(pdx_ctx *c, int *bond);
/* How K follows bond failure (FC 2 in the quasi-static, explicit and Newmark drivers): on (the
 * default), the rows of K of the families whose bonds changed are recomputed on the last whole
 * assembly (the family kernel for those families, then their node pairs with element terms),
 * which gives the same K bit for bit; off, every change assembles the whole K again, as the
 * reference rebuilds its system every step.  No counterpart in the reference's interface. */
int pdx_set_incremental_bonds(pdx_ctx *c, int on);
/* calLocalDamage (pdsolve.cpp:7338-7373): phi [n_node] = 1 - intact / all bonds of the node's
 * family (0 for nodes without a family), from the context's bond states */
int pdx_local_damage(pdx_ctx *c, double *phi);
/* Nodal stresses, pdsolve::calGlobalNodeStresses (pdsolve.cpp:9540-9604) with
 * calPDNodeStresses (:9081-9166) and calFEMNodeStresses_EXP (:9168-9335), as the
 * drivers call it before every result output (e.g. :6062).  u_all: DOF values
 * [n_node*dim] (the reference's pdDof values), or NULL for the DOF values of the last stored
 * explicit step (free DOFs u_{n+1}, prescribed DOFs at the values that step's force used, as
 * storeDisplacementResult leaves them).  sigma [n_node*6]: xx, yy, zz, xy, yz, zx (2D: xx, yy,
 * zz, xy). */
int pdx_node_stresses(pdx_ctx *c, const double *u_all, double *sigma);
/* The same recovery without allocating.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  u_dev [n_node*dim] and sigma_dev [n_node*6] are device pointers
 * owned by the caller; the element corner stresses, family stresses and the status word live in
 * a fixed 96 MiB device array of the library (one per GPU, held until the process ends):
 * models needing more (8*6 doubles per FE element + 6 per family + 1) get PDX_ERR_LIMIT.  K must
 * have been assembled on the context (the recovery reads the bond strain weights the assembly
 * leaves).  The launches go to the context's stream and are not waited for; no memset, host
 * copy or synchronisation, so they may be captured.  Calls on one GPU share the array: calls
 * that overlap in time on one GPU (two streams, two threads) race. */
int pdx_node_stresses_noalloc(pdx_ctx *c, const double *u_dev, double *sigma_dev);
/* the same on the device only, timed (ms of one recovery, inputs resident) */
int pdx_node_stresses_timed(pdx_ctx *c, const double *u_all, double *ms);
/* Newmark (acceleration form), pdfemDynamicNewmarkSolver_CSRformat (pdsolve.cpp:6092-6290): per
 * step the predictor U_N (ramped prescribed values advanced by dt * velocity), F = F_ext - K U_N,
 * (M + beta dt^2 K) a_{n+1} = F by the Krylov solve of o (warm-started from a_n), the update, then
 * the BCs advance; with FC 2 in 2D the nodal stresses and the bond failure follow (:6258-6261,
 * pdx_step_log) and K is assembled again when bonds broke.  u, v, a: [neq] (may be NULL).
 * PDX_PREC_LEARNED predicts the smoother of each K_eff with the network of pdx_gnn_attach.  st:
 * iterations and refinements summed over the steps, the largest true residual, the preconditioner
 * and precision of the last step's solve, the most attempts any step took. */
int pdx_newmark_run(pdx_ctx *c, int nsteps, const pdx_solve_opts *o, double *u, double *v, double *a,
                    pdx_solve_stats *st);

/* ---- learned smoother of the two-level preconditioner ------------------
 * No counterpart in the reference (which factorises with PARDISO,
 * calMatrixOperations.cpp:378-498): a graph network over the node graph of K
 * predicts a bounded correction of the node-block Jacobi smoother,
 * M_a = (I + 0.5 tanh(o_a)) (K_aa)^-1, used by PDX_PREC_LEARNED.  Training
 * minimises ||f - M K f||^2 / ||f||^2 over coarse-corrected Gaussian probes
 * f = (I - Q K) z, on the GPU. */
typedef struct pdx_gnn pdx_gnn;
int pdx_gnn_create(int dim, int device, unsigned long long seed, pdx_gnn **out);
void pdx_gnn_destroy(pdx_gnn *g);
long long pdx_gnn_num_params(const pdx_gnn *g);
int pdx_gnn_get_params(const pdx_gnn *g, float *w);
int pdx_gnn_set_params(pdx_gnn *g, const float *w); /* also resets the Adam state */
/* steps Adam steps on n_probes probes of the context's current K; loss: [steps+1] or NULL */
int pdx_gnn_train(pdx_ctx *c, pdx_gnn *g, int n_probes, unsigned long long probe_seed, int coarse_agg, int steps,
                  float lr, double *loss);
/* loss and gradient (flat, pdx_gnn_num_params floats) of the current weights */
int pdx_gnn_loss_grad(pdx_ctx *c, pdx_gnn *g, int n_probes, unsigned long long probe_seed, int coarse_agg,
                      double *loss, float *grad);
/* the network inputs: X [nn*8], c [nblk], A_aa [nn*d*d], B [nn*d*d], f, y [n_probes*nn*d]; any may be NULL */
int pdx_gnn_inputs(pdx_ctx *c, int n_probes, unsigned long long probe_seed, int coarse_agg, float *X, float *cw,
                   float *Aaa, float *B, float *f, float *y);
/* predict the smoother for the context's current operator (pdx_gnn_operator; used by
 * PDX_PREC_LEARNED on that operator) */
int pdx_gnn_apply(pdx_ctx *c, const pdx_gnn *g);
/* The operator the network's inputs, training probes and pdx_gnn_apply / pdx_gnn_infer use:
 * mode 0 (default) the stiffness K of the static and quasi-static solves, mode 1 the Newmark
 * operator K_eff = M + beta dt^2 K of pdfemDynamicNewmarkSolver_CSRformat (pdsolve.cpp:6092-6290),
 * formed on the device for the current K.  A smoother serves only the operator it was predicted
 * for; a solve on another operator reports that none is there (or falls back). */
int pdx_gnn_operator(pdx_ctx *c, int mode);
/* Network a learned Newmark run (pdx_newmark_run with opts.precond = PDX_PREC_LEARNED) predicts
 * its smoother with, for each K_eff it forms (again after bonds break); NULL detaches.  The
 * caller keeps the network alive while it is attached.  No counterpart in the reference. */
int pdx_gnn_attach(pdx_ctx *c, const pdx_gnn *g);
/* The same prediction with FP32 (bf16 = 0) or BF16-activation MFMA (bf16 != 0,
 * v_mfma_f32_16x16x32_bf16, FP32 accumulation) inference; M [n_node*d*d] (or NULL) receives the
 * smoother, ms (or NULL) the inference time.  Also makes it the context's learned smoother. */
int pdx_gnn_infer(pdx_ctx *c, const pdx_gnn *g, int bf16, double *M, double *ms);

/* ---- batch of PD families (a node cloud, every node a family) ------------
 * The family kernel of the reference -- inflFunc, shapTens2D/3D, vec_gd2D/3D
 * with PLUSolve, matG2D/3D, matH2D/3D (pdsolve.cpp:1929-2416) -- for a whole
 * cloud with per-node horizons.  Members by the rule of initialPDN_Fami
 * (pdsolve.cpp:1442-1568): j joins k iff |x_j - x_k| < delta_k or < delta_j
 * (self excluded), self first, then in cell-scan order.  Results stay on the
 * device: A (FP64, p x p per family, p = 5 / 9), the lower Cholesky factor L of
 * its symmetric part S (A = S diag(s), s = 1/2 on the squared terms, S = L L^T),
 * and H (FP32, d x d per member; member 0 holds minus the sum of the others). */
typedef struct pdx_fbatch pdx_fbatch;
int pdx_fb_create(int device, int dim, long long n, const double *xyz /*[n*3]*/, const double *vol /*[n]*/,
                  const double *delta /*[n]*/, double nlf, double lambda, double mu, pdx_fbatch **out);
void pdx_fb_destroy(pdx_fbatch *b);
int pdx_fb_sizes(const pdx_fbatch *b, long long *n_fam, long long *n_bond, int *max_family);
int pdx_fb_families(const pdx_fbatch *b, long long *off /*[n+1]*/, int *mem /*[n_bond]*/);
/* variant -1: the fastest launch form (every form timed on the first call, one family per block
 * included); 0: one family per 64-thread block; 1..6: persistent, (waves per block, families
 * per chunk) = (4, 4), (2, 4), (1, 4), (4, 16), (2, 16), (1, 16).  ms: kernel time */
#define PDX_FB_VARIANTS 7
int pdx_fb_run(pdx_fbatch *b, int variant, float *ms);
int pdx_fb_tuning(const pdx_fbatch *b, int *best_variant, float *ms /*[PDX_FB_VARIANTS]*/);
int pdx_fb_results(const pdx_fbatch *b, long long first, long long count, double *A, double *L, float *H);

/* PARDISO seam: solve A x = b for a zero-based int64 CSR (n, ia, ja, a). */
int pdx_solve_csr(int device, long long n, const long long *ia, const long long *ja, const double *a,
                  const double *b, double *x, const pdx_solve_opts *o, pdx_solve_stats *st);

/* timing of the last pdx_explicit_steps call measured with HIP events on the
 * context stream: total ms and per-step ms */
int pdx_last_timing(pdx_ctx *c, double *total_ms, double *per_step_ms);
/* the explicit-step kernel form the tuner chose (lanes per short row, non-temporal K loads,
 * XCD-contiguous block order, row-local value order of K, column pattern classes: a row's block
 * columns read as its index plus the offsets of its class, shared by the rows of one stencil;
 * mirrored blocks: a lower block that is the exact transpose of its upper partner (both summed
 * from element matrices only) is read from the partner's storage; alternating: every other step
 * visits the rows last to first, so a step begins on the data the previous one read last (forms
 * 20-23); all 0 before the first step)
 * and its timings of the PDX_EXPLICIT_FORMS forms (ms per step) */
#define PDX_EXPLICIT_FORMS 24
int pdx_explicit_tuning(pdx_ctx *c, int precision, int *lanes, int *nontemporal, int *xcd_order, int *row_soa,
                        int *col_classes, int *mirrored, int *alternating, int *form,
                        float *ms /*[PDX_EXPLICIT_FORMS]*/);
/* use explicit-step form `form` (0 .. PDX_EXPLICIT_FORMS-1) from now on instead of the tuner's
 * choice, or -1 to tune again at the next step; forms with the same lanes per row compute the
 * same bits, the lanes per row change the order of the row sum */
int pdx_explicit_set_form(pdx_ctx *c, int precision, int form);

#ifdef __cplusplus
}
#endif
#endif
