// Host-side data model of the PDLSM-FEM hot path: mesh, DOF numbering,
// PD/FE classification, PD-region boundary entities, families and the
// stiffness pattern.  This is the drop-in counterpart of the reference's
// datModel + pdsolve set-up (SourceFiles/Solver/datModel.cpp,
// SourceFiles/Solver/pdsolve.cpp:114-151, :959-1121); the device context is
// built from it.
#pragma once
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdx.h"

namespace pdx {

struct EbcSet {
    std::vector<int> nid;  // 0-based nodes
    int dof = 0;
    double val = 0, vel = 0;
    bool vary = false;
};

struct NbcSet {
    std::vector<int> con;  // [n_ele * tn] 0-based nodes
    double val = 0, vel = 0;
    bool vary = false;
};

struct Model {
    // ---- raw mesh ----
    int dim = 3, pro_type = -1;
    double E = 0, nu = 0, rho = 0;
    int nn = 0, ne = 0;
    std::vector<double> x;    // [nn*3]
    std::vector<int> econ;    // [ne*8] 0-based
    std::vector<int> nen;     // [ne]
    std::vector<int> algo;    // [ne] 1 PD, 2 FE (after crack conversion)
    std::vector<EbcSet> ebc;
    std::vector<NbcSet> nbc;
    std::vector<double> crack;  // [n_crack*9]
    double fac = 3, nlf = 0.33, me = 0, dt = 1, beta = 0.25, gamma = 0.5;
    std::string title = "pdlsm";  // first token of the mesh file (the VTK title)
    double sig_ult = 0, kic = 0;  // material strength and toughness of the mesh file
    int fc = 0, topk = 0;         // failure criterion (FC) and bonds per step (TOPK) of the command file
    // run settings of the command file (datModel defaults, datModel.cpp:17-56)
    int solver_flag = -1;          // SOLVER: 0 dynamic, 1 static, 2 quasi-static
    int num_steps = 1, save_freq = 1;  // SETSOLVING tokens 2 and 3
    bool newmark = false;          // NEWMARK ON|OFF (cb_Newmark)
    bool lumped = false;           // LUMPMASS ON|OFF (cb_lumpedMass)
    bool vtk_binary = true;        // VTKFORMAT BINARY|ASCII (cb_vtkBinary)
    std::vector<int> rf_ids;       // RF: EBC set ids of the reaction-force output

    // ---- DOFs ----
    std::vector<int> eq;       // [nn*3], -1 prescribed / unused
    std::vector<double> dofval;  // [nn*3] current DOF values
    std::vector<int> presc_set;  // [nn*3] EBC set of a prescribed DOF, -1 otherwise
    int neq = 0;

    // ---- geometry ----
    std::vector<double> ev, nv;  // element / node volumes
    double minDelta = 0, maxDelta = 0, bsize = 0;
    double lbc[3] = {0, 0, 0}, rtc[3] = {0, 0, 0};
    int nb[3] = {1, 1, 1};
    std::vector<int> cell_node_off, cell_node;  // nodes per cell (node order)
    std::vector<int> cell_ele_off, cell_ele;    // elements per cell (element order)

    // ---- PD / FE split ----
    std::vector<int> pde, fe;        // element lists (reference order)
    std::vector<int> crac_ele;
    std::vector<int> pdbe;           // [npdbe*4] 0-based; 2D uses [0..1]
    int npdbe = 0;
    std::vector<int> ntype;          // 0 FE, 1 interface, 2 PD
    std::vector<int> famid;          // family index or -1

    // ---- families ----
    int nfam = 0, max_family = 0;
    std::vector<int> fam_node;       // node of family k
    std::vector<int> fam_off;        // [nfam+1]
    std::vector<int> fam_mem;        // 0-based node ids, self first
    std::vector<int> bond;           // 1 intact, 0 broken
    std::vector<double> fam_v;       // cd_v: PD-element volume share
    std::vector<double> fam_delta;   // horizon

    // ---- node graph (pattern at node level) and CSR ----
    std::vector<int> adj_off, adj;   // sorted unique neighbour nodes per node
    std::vector<int64_t> ia;         // K_ff CSR row offsets, zero based
    int64_t nnz = 0;
    // K_ff CSR column indices (sorted per row), built on first use and kept
    const std::vector<int64_t> &csr_ja() const;
    mutable std::vector<int64_t> ja_store;
    mutable bool ja_built = false;
    mutable std::mutex ja_mu;

    // ---- material ----
    double D[36] = {0};
    double lambda = 0, mu = 0;

    int setup(std::string &err);
};

int model_from_desc(const pdx_mesh_desc &d, Model &m, std::string &err);
int model_read_files(const std::string &cmd, Model &m, int &solver_flag, int &num_steps, std::string &err);

// consistent-mass pattern of the explicit driver (setCSRIndexes_gloMassMat, pdsolve.cpp:5674-5817):
// zero-based int64 CSR over the free DOFs, columns sorted per row
void mass_pattern(const Model &m, std::vector<int64_t> &ia, std::vector<int64_t> &ja);

// small numerics shared with the device set-up
double det33(const double m[3][3]);

// updateBondstate (pdsolve.cpp:8910-9004): breaks the intact bonds (in `bond`, laid out like
// Model::bond) that cross the crack (2D segment / 3D triangle, 9 doubles); returns how many
int break_bonds_across(const Model &m, const double *crack, std::vector<int> &bond);

// result file in the reference's VTK layout (vtk.cpp); null arrays write zeros
int write_vtk(const Model &m, const std::string &path, const std::string &title, const double *u,
              const double *sigma, const double *phi, bool binary, std::string &err);
// the PD-only (part 0) or FE-only (part 1) ASCII result file
int write_vtk_part(const Model &m, const std::string &path, const std::string &title, const double *u,
                   const double *sigma, const double *phi, int part, std::string &err);
// the .disp (kind 0) or .sig (kind 1) node table
int write_node_table(const Model &m, const std::string &path, const double *u, const double *sigma, int kind,
                     std::string &err);

}  // namespace pdx
