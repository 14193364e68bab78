// Device context construction (upload of the host model in the HBM layout of
// device.hpp) and the extern "C" entry points of libpdx.so (include/pdx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "../../include/pdx.h"
#include "device.hpp"
#include "setup.hpp"

namespace pdx {

int dev_explicit_steps(Ctx &c, int nsteps, int prec, int use_graph, bool lumped,
                       const std::function<int(std::string &)> &after_step, std::string &err);
int dev_explicit_dofvals(Ctx &c, const double **out, std::string &err);
int dev_newmark(Ctx &c, int nsteps, const pdx_solve_opts &o, std::vector<double> &dofval, double *u_h, double *v_h,
                double *a_h, pdx_solve_stats &stats, const std::function<int(const double *, std::string &)> &after_step,
                std::string &err);
int dev_explicit_state(Ctx &c, double *u_out, double *u_prev, std::string &err);
int dev_max_eig(Ctx &c, int iters, double *lmax, std::string &err);
int stress_scratch_init(int device, std::string &err);
int dev_node_stresses_noalloc(Ctx &c, const double *u_dev, double *sigma_dev, std::string &err);

int dev_solve(Ctx &c, const pdx_solve_opts &o, const double *b_dev, double *x_dev, pdx_solve_stats &st,
              std::string &err);

template <class T>
static int up(Ctx &c, T *&dst, const std::vector<T> &src, std::string &err) {
    size_t n = std::max<size_t>(src.size(), 1);
    PDX_HIP(hipMalloc((void **)&dst, n * sizeof(T)));
    c.owned.push_back(dst);
    if (!src.empty()) PDX_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return PDX_OK;
}

template <class T>
static int alloc(Ctx &c, T *&dst, size_t n, std::string &err) {
    PDX_HIP(hipMalloc((void **)&dst, std::max<size_t>(n, 1) * sizeof(T)));
    c.owned.push_back(dst);
    PDX_HIP(hipMemset(dst, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return PDX_OK;
}

// statement macros (do/while, so an else after one binds to the caller's if)
#define UP(dst, src)                                       \
    do {                                                   \
        if ((rc = up(*c, dst, src, err))) return rc;       \
    } while (0)
#define AL(dst, n)                                         \
    do {                                                   \
        if ((rc = alloc(*c, dst, n, err))) return rc;      \
    } while (0)

static int find_block(const Model &m, int row_node, int col_node) {
    auto b = m.adj.begin() + m.adj_off[row_node], e = m.adj.begin() + m.adj_off[row_node + 1];
    auto it = std::lower_bound(b, e, col_node);
    if (it == e || *it != col_node) return -1;
    return int(it - m.adj.begin());
}

int ctx_build(const Model &m, int device, Ctx *&out, std::string &err) {
    int rc = PDX_OK;
    PDX_HIP(hipSetDevice(device));
    Ctx *c = new Ctx();
    out = c;
    c->device = device;
    c->host = &m;
    PDX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const int D = m.dim;
    c->dim = D;
    c->dd = D * D;
    c->nn = m.nn;
    c->ne = m.ne;
    c->neq = m.neq;
    c->nfam = m.nfam;
    c->max_family = std::max(m.max_family, 1);
    c->npdbe = m.npdbe;
    c->nblk = m.adj_off[m.nn];
    c->nnz = m.nnz;
    c->lambda = m.lambda;
    c->mu = m.mu;
    c->nlf = m.nlf;
    c->rho = m.rho;
    c->dt = m.dt;
    c->beta = m.beta;
    c->gamma = m.gamma;

    UP(c->x, m.x);
    UP(c->nv, m.nv);
    UP(c->ev, m.ev);
    UP(c->Dm, std::vector<double>(m.D, m.D + 36));
    UP(c->econ, m.econ);
    UP(c->nen, m.nen);
    UP(c->eq, m.eq);
    UP(c->presc_set, m.presc_set);
    UP(c->adj_off, m.adj_off);
    UP(c->adj, m.adj);
    UP(c->ia, m.ia);

    // families + scatter map (block index of every member in the family node's row)
    std::vector<int> fam_blk(m.fam_mem.size());
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int f = 0; f < m.nfam; f++)
        for (int q = m.fam_off[f]; q < m.fam_off[f + 1]; q++) {
            const int b = find_block(m, m.fam_node[f], m.fam_mem[q]);
            if (b < 0) bad = 1;
            else fam_blk[q] = b;
        }
    if (bad) { err = "Error: can not build matrix in CSR format (family member)"; return PDX_ERR_PATTERN; }
    UP(c->fam_node, m.fam_node);
    UP(c->fam_off, m.fam_off);
    UP(c->fam_mem, m.fam_mem);
    UP(c->bond, m.bond);
    UP(c->fam_blk, fam_blk);
    // K blocks the family kernel stores (every member block of every family, broken bonds included)
    std::vector<unsigned char> pd_bits(((size_t)c->nblk + 7) / 8, 0);
    for (int b : fam_blk) pd_bits[(size_t)b >> 3] |= (unsigned char)(1u << (b & 7));
    UP(c->pd_bits, pd_bits);
    // mirrored storage for the explicit step: a lower block (r, c) none of whose two orientations
    // a family stores is a sum of element blocks only, over the same elements in the same order
    // as its upper partner (c, r) with every element block transposed, so it equals the partner's
    // transpose exactly and is read from the partner's storage
    {
        auto pd = [&](long long b) { return (pd_bits[(size_t)b >> 3] >> (b & 7)) & 1; };
        std::vector<int> ksrc((size_t)std::max<long long>(c->nblk, 1)), partner((size_t)c->nblk, -1);
        std::vector<long long> nown((size_t)m.nn + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
        for (int r = 0; r < m.nn; r++) {
            long long own = 0;
            for (int b = m.adj_off[r]; b < m.adj_off[r + 1]; b++) {
                const int col = m.adj[b];
                if (col < r && !pd(b)) {
                    const int tb = find_block(m, col, r);
                    if (tb >= 0 && !pd(tb)) { partner[b] = tb; continue; }
                }
                own++;
            }
            nown[r + 1] = own;
        }
        for (int r = 0; r < m.nn; r++) nown[r + 1] += nown[r];
        // own blocks numbered in block order
        std::vector<int> own_slot((size_t)nown[m.nn]);
#pragma omp parallel for schedule(dynamic, 256)
        for (int r = 0; r < m.nn; r++) {
            long long s = nown[r];
            for (int b = m.adj_off[r]; b < m.adj_off[r + 1]; b++)
                if (partner[b] < 0) {
                    ksrc[b] = int(s);
                    own_slot[(size_t)s++] = b;
                }
        }
#pragma omp parallel for schedule(static)
        for (long long b = 0; b < c->nblk; b++)
            if (partner[b] >= 0) ksrc[b] = ~ksrc[partner[b]];  // partners are upper blocks, always own
        c->n_own = (long long)own_slot.size();
        UP(c->ksrc, ksrc);
        UP(c->kown_slot, own_slot);
    }
    UP(c->fam_v, m.fam_v);
    UP(c->fam_delta, m.fam_delta);
    // PD-boundary entities incident to each family's node, in entity order
    {
        std::vector<std::vector<int>> inc(m.nfam);
        const int per = D == 2 ? 2 : 4;
        for (int p = 0; p < m.npdbe; p++) {
            const int *con = &m.pdbe[p * 4];
            int cnt = per;
            if (D == 3 && con[3] == con[2]) cnt = 3;
            for (int a = 0; a < cnt; a++) {
                int f = m.famid[con[a]];
                if (f < 0) { err = "Warning: This node of PD domain boundary is not PD node!"; return PDX_ERR_MESH; }
                inc[f].push_back(p);
            }
        }
        std::vector<int> off(m.nfam + 1, 0), lst;
        for (int f = 0; f < m.nfam; f++) {
            off[f + 1] = off[f] + int(inc[f].size());
            lst.insert(lst.end(), inc[f].begin(), inc[f].end());
        }
        UP(c->fp_off, off);
        UP(c->fp_list, lst);
        c->max_pdbe_per_family = 0;
        for (int f = 0; f < m.nfam; f++) c->max_pdbe_per_family = std::max(c->max_pdbe_per_family, off[f + 1] - off[f]);
        std::vector<int> pd(m.pdbe);
        UP(c->pdbe, pd);
    }
    // FE element rows: node -> (fe position, local index) in element order; block map
    {
        c->nfe = int(m.fe.size());
        std::vector<int> fe_blk((size_t)c->nfe * 64, -1), cnt(m.nn + 1, 0);
        int bad_fe = 0, rep_fe = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_fe, rep_fe)
        for (int t = 0; t < c->nfe; t++) {
            const int e = m.fe[t];
            for (int a = 0; a < m.nen[e]; a++)
                for (int b = 0; b < m.nen[e]; b++) {
                    if (a != b && m.econ[(size_t)e * 8 + a] == m.econ[(size_t)e * 8 + b]) rep_fe = 1;
                    const int bi = find_block(m, m.econ[(size_t)e * 8 + a], m.econ[(size_t)e * 8 + b]);
                    if (bi < 0) bad_fe = 1;
                    else fe_blk[(size_t)t * 64 + a * 8 + b] = bi;
                }
        }
        if (bad_fe) { err = "Error: can not build matrix in CSR format (element)"; return PDX_ERR_PATTERN; }
        if (rep_fe) { err = "an FE element lists one node twice"; return PDX_ERR_MESH; }
        for (int t = 0; t < c->nfe; t++) {
            const int e = m.fe[t];
            for (int a = 0; a < m.nen[e]; a++) cnt[m.econ[(size_t)e * 8 + a] + 1]++;
        }
        for (int k = 0; k < m.nn; k++) cnt[k + 1] += cnt[k];
        std::vector<int> item(cnt[m.nn]), fill(cnt.begin(), cnt.end() - 1);
        for (int t = 0; t < c->nfe; t++) {
            int e = m.fe[t];
            for (int a = 0; a < m.nen[e]; a++) item[fill[m.econ[(size_t)e * 8 + a]]++] = t * 8 + a;
        }
        UP(c->fe_elem, m.fe);
        UP(c->fe_blk, fe_blk);
        std::vector<int> la, lb;
        for (int t = 0; t < c->nfe; t++) (m.nen[m.fe[t]] == (D == 3 ? 8 : 4) ? la : lb).push_back(t);
        c->n_fe_a = int(la.size());
        c->n_fe_b = int(lb.size());
        UP(c->fe_list_a, la);
        UP(c->fe_list_b, lb);
        UP(c->nfe_off, cnt);
        UP(c->nfe_item, item);
        // per K block, its FE contributions in element order, for the block-owned gather
        {
            // the blocks a node's element rows reach lie in that node's row, so the nodes can be
            // processed in parallel
            std::vector<int> fcnt((size_t)c->nblk + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
            for (int k = 0; k < m.nn; k++)
                for (int q = cnt[k]; q < cnt[k + 1]; q++) {
                    const int t = item[q] >> 3, a = item[q] & 7;
                    for (int b = 0; b < m.nen[m.fe[t]]; b++) fcnt[(size_t)fe_blk[(size_t)t * 64 + a * 8 + b] + 1]++;
                }
            for (long long q = 0; q < c->nblk; q++) fcnt[q + 1] += fcnt[q];
            // contributions as staged block indices (FE-list position, local row node, local column node)
            const int NB = m.dim == 3 ? 8 : 4, NP = NB * (NB + 1) / 2;
            if ((long long)m.fe.size() * NP * 2 > 0x7fffffffLL) {
                err = "too many FE elements for 32-bit staged block indices";
                return PDX_ERR_LIMIT;
            }
            std::vector<int> fec(std::max(fcnt[c->nblk], 1)), ffill(fcnt.begin(), fcnt.end() - 1);
#pragma omp parallel for schedule(dynamic, 256)
            for (int k = 0; k < m.nn; k++)
                for (int q = cnt[k]; q < cnt[k + 1]; q++) {
                    const int t = item[q] >> 3, a = item[q] & 7;
                    for (int b = 0; b < m.nen[m.fe[t]]; b++) {
                        // upper staged block (min, max) of the node pair; transposed when a > b
                        const int lo = std::min(a, b), hi = std::max(a, b);
                        const int pair = lo * NB - lo * (lo - 1) / 2 + (hi - lo);
                        fec[ffill[fe_blk[(size_t)t * 64 + a * 8 + b]]++] = ((t * NP + pair) << 1) | (a > b ? 1 : 0);
                    }
                }
            // the gather's node pairs (every upper block with FE terms, with its lower partner) and
            // the staging slots: the element blocks of one pair are staged consecutively, in
            // element order, each in the pair's (r, c) orientation
            {
                std::vector<int> npr((size_t)m.nn + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
                for (int r = 0; r < m.nn; r++) {
                    int n = 0;
                    for (int b = m.adj_off[r]; b < m.adj_off[r + 1]; b++) n += m.adj[b] >= r && fcnt[b + 1] > fcnt[b];
                    npr[r + 1] = n;
                }
                for (int r = 0; r < m.nn; r++) npr[r + 1] += npr[r];
                std::vector<int> pu((size_t)npr[m.nn]), pl((size_t)npr[m.nn]), prow((size_t)npr[m.nn]);
                int bad_pair = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad_pair)
                for (int r = 0; r < m.nn; r++) {
                    int q = npr[r];
                    for (int b = m.adj_off[r]; b < m.adj_off[r + 1]; b++) {
                        const int col = m.adj[b];
                        if (col < r || fcnt[b + 1] == fcnt[b]) continue;
                        const int l = col > r ? find_block(m, col, r) : -1;
                        if (col > r && l < 0) bad_pair = 1;
                        pu[q] = b;
                        prow[q] = r;
                        pl[q++] = l;
                    }
                }
                if (bad_pair) { err = "Error: can not build matrix in CSR format (element pair)"; return PDX_ERR_PATTERN; }
                const int ntask = npr[m.nn];
                // pairs without a PD part in either block first (the gather adds them while the
                // family kernel still runs), then the others, each part in row order
                {
                    auto pdb = [&](int b) { return (pd_bits[(size_t)b >> 3] >> (b & 7)) & 1; };
                    std::vector<int> pu2, pl2, prow_pd;
                    pu2.reserve(ntask);
                    pl2.reserve(ntask);
                    for (int pass = 0; pass < 2; pass++) {
                        for (int q = 0; q < ntask; q++) {
                            const bool with_pd = pdb(pu[q]) || (pl[q] >= 0 && pdb(pl[q]));
                            if (with_pd == (pass == 1)) {
                                pu2.push_back(pu[q]);
                                pl2.push_back(pl[q]);
                                if (pass == 1) prow_pd.push_back(prow[q]);
                            }
                        }
                        if (pass == 0) c->n_fe_pair_fe = int(pu2.size());
                    }
                    pu.swap(pu2);
                    pl.swap(pl2);
                    UP(c->fe_pair_row, prow_pd);
                }
                std::vector<long long> poff((size_t)ntask + 1, 0);
                for (int q = 0; q < ntask; q++) poff[q + 1] = poff[q] + (fcnt[pu[q] + 1] - fcnt[pu[q]]);
                if (poff[ntask] >= (1LL << 30)) {
                    err = "too many FE element blocks for 32-bit staging slots";
                    return PDX_ERR_LIMIT;
                }
                std::vector<int> slot((size_t)c->nfe * NP, -1), po((size_t)ntask + 1);
                for (int q = 0; q <= ntask; q++) po[q] = int(poff[q]);
                // every element pair belongs to one upper block (the element's nodes are distinct)
#pragma omp parallel for schedule(dynamic, 1024)
                for (int q = 0; q < ntask; q++)
                    for (int k = fcnt[pu[q]]; k < fcnt[pu[q] + 1]; k++)
                        slot[(size_t)(fec[k] >> 1)] = ((po[q] + (k - fcnt[pu[q]])) << 1) | (fec[k] & 1);
                c->n_fe_pair = ntask;
                c->n_stage = poff[ntask];
                UP(c->fe_pair_up, pu);
                UP(c->fe_pair_lo, pl);
                UP(c->pair_off, po);
                UP(c->fe_slot, slot);
            }
            // a block neither stored by a family nor gathered from elements keeps the fill's zero;
            // without such blocks the assembly needs no fill of K
            long long bare = 0;
#pragma omp parallel for schedule(static) reduction(+ : bare)
            for (long long q = 0; q < c->nblk; q++)
                if (!((pd_bits[(size_t)q >> 3] >> (q & 7)) & 1) && fcnt[q + 1] == fcnt[q]) bare++;
            c->k_needs_fill = bare > 0;
        }
        // all elements (lumped mass)
        std::vector<int> acnt(m.nn + 1, 0);
        for (int e = 0; e < m.ne; e++)
            for (int a = 0; a < m.nen[e]; a++) acnt[m.econ[(size_t)e * 8 + a] + 1]++;
        for (int k = 0; k < m.nn; k++) acnt[k + 1] += acnt[k];
        std::vector<int> aitem(acnt[m.nn]), afill(acnt.begin(), acnt.end() - 1);
        for (int e = 0; e < m.ne; e++)
            for (int a = 0; a < m.nen[e]; a++) aitem[afill[m.econ[(size_t)e * 8 + a]]++] = e;
        UP(c->nall_off, acnt);
        UP(c->nall_ele, aitem);
    }
    // rows: equation -> node*3+dof, and node rows binned by length
    {
        std::vector<int> rn(std::max(m.neq, 1), 0);
        for (int k = 0; k < m.nn; k++)
            for (int i = 0; i < D; i++)
                if (m.eq[(size_t)k * 3 + i] >= 0) rn[m.eq[(size_t)k * 3 + i]] = k * 3 + i;
        UP(c->row_node, rn);
        std::vector<int> rs, rl;
        for (int k = 0; k < m.nn; k++) (m.adj_off[k + 1] - m.adj_off[k] > 48 ? rl : rs).push_back(k);
        c->n_short = int(rs.size());
        c->n_long = int(rl.size());
        UP(c->rows_short, rs);
        UP(c->rows_long, rl);
    }
    // column pattern classes: rows whose block columns have the same offsets from the row index
    // (the stencil of a structured mesh, jittered or not) share one list of offsets
    {
        // a hash of each row's offsets in parallel, then classes numbered in order of first
        // appearance, rows with equal hashes compared offset by offset
        std::vector<uint64_t> hsh(m.nn);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < m.nn; r++) {
            uint64_t h = 1469598103934665603ull ^ uint64_t(m.adj_off[r + 1] - m.adj_off[r]);
            for (int b = m.adj_off[r]; b < m.adj_off[r + 1]; b++) h = (h ^ uint64_t(uint32_t(m.adj[b] - r))) * 1099511628211ull;
            hsh[r] = h;
        }
        std::unordered_map<uint64_t, std::vector<int>> ids;
        std::vector<int> cls(m.nn), coff(1, 0), crel;
        int ncls = 0;
        for (int r = 0; r < m.nn; r++) {
            const int b0 = m.adj_off[r], len = m.adj_off[r + 1] - b0;
            std::vector<int> &cand = ids[hsh[r]];
            int found = -1;
            for (int id : cand) {
                if (coff[id + 1] - coff[id] != len) continue;
                bool same = true;
                for (int j = 0; j < len && same; j++) same = crel[(size_t)coff[id] + j] == m.adj[b0 + j] - r;
                if (same) { found = id; break; }
            }
            if (found < 0) {
                found = ncls++;
                cand.push_back(found);
                for (int j = 0; j < len; j++) crel.push_back(m.adj[b0 + j] - r);
                coff.push_back(int(crel.size()));
            }
            cls[r] = found;
        }
        c->n_cls = ncls;
        UP(c->cls, cls);
        UP(c->cls_off, coff);
        UP(c->cls_rel, crel);
    }
    // natural BCs
    {
        c->n_nbc = int(m.nbc.size());
        const int tn = D == 3 ? 4 : 2;
        std::vector<int> con, set;
        for (int s = 0; s < c->n_nbc; s++) {
            con.insert(con.end(), m.nbc[s].con.begin(), m.nbc[s].con.end());
            for (size_t e = 0; e < m.nbc[s].con.size() / tn; e++) set.push_back(s);
            c->nbc_val.push_back(m.nbc[s].val);
            c->nbc_vel.push_back(m.nbc[s].vel);
            c->nbc_vary.push_back(m.nbc[s].vary);
        }
        c->n_nbc_ele = int(set.size());
        std::vector<int> cnt(m.nn + 1, 0);
        for (int e = 0; e < c->n_nbc_ele; e++)
            for (int a = 0; a < tn; a++) cnt[con[(size_t)e * tn + a] + 1]++;
        for (int k = 0; k < m.nn; k++) cnt[k + 1] += cnt[k];
        std::vector<int> item(cnt[m.nn]), fill(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < c->n_nbc_ele; e++)
            for (int a = 0; a < tn; a++) item[fill[con[(size_t)e * tn + a]]++] = e * tn + a;
        UP(c->nbc_con, con);
        UP(c->nbc_set, set);
        UP(c->nbcn_off, cnt);
        UP(c->nbcn_item, item);
        AL(c->nbc_fe, (size_t)c->n_nbc_ele * 12);
        AL(c->f_unit, (size_t)std::max(c->n_nbc, 1) * m.nn * D);
        AL(c->nbcv, 2 * (size_t)std::max(c->n_nbc, 1));
        AL(c->d_nbc_vel, (size_t)std::max(c->n_nbc, 1));
        for (auto &s : m.ebc) {
            c->ebc_val.push_back(s.val);
            c->ebc_vel.push_back(s.vel);
            c->ebc_vary.push_back(s.vary);
        }
        AL(c->d_ebc_vel, std::max<size_t>(m.ebc.size(), 1));
    }
    AL(c->K, (size_t)c->dd * c->nblk);
    AL(c->mlump, (size_t)m.nn * D);
    for (int i = 0; i < 3; i++) AL(c->U[i], (size_t)m.nn * D);
    // the zero fills above run on the null stream: finish them before any work on c->stream
    PDX_HIP(hipDeviceSynchronize());
    if ((rc = dev_nbc_unit(*c, err))) return rc;
    PDX_HIP(hipStreamSynchronize(c->stream));
    // node -> family id of the stress recovery, and the recovery's allocation-free scratch on
    // this device (its address is taken once per device, with the device current)
    {
        std::vector<int> fid(m.famid.begin(), m.famid.end());
        UP(c->stress_famid, fid);
    }
    if ((rc = stress_scratch_init(device, err))) return rc;
    return PDX_OK;
}

void ctx_free(Ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &g : c->graph)
        if (g) hipGraphExecDestroy(g);
    for (void *p : c->owned) hipFree(p);
    if (c->K32) hipFree(c->K32);
    if (c->nlm_F) hipFree(c->nlm_F);
    if (c->nlm_a) hipFree(c->nlm_a);
    if (c->dofs_tmp) hipFree(c->dofs_tmp);
    if (c->K32s) hipFree(c->K32s);
    if (c->K64s) hipFree(c->K64s);
    if (c->Mk) hipFree(c->Mk);
    if (c->Ke_stage) hipFree(c->Ke_stage);
    if (c->d_flist) hipFree(c->d_flist);
    if (c->d_rowflag) hipFree(c->d_rowflag);
    if (c->red) hipFree(c->red);
    if (c->learned_Minv) hipFree(c->learned_Minv);
    if (c->gnn_keff) hipFree(c->gnn_keff);
    if (c->fam_c) hipFree(c->fam_c);
    if (c->Kown32) hipFree(c->Kown32);
    if (c->Kown64) hipFree(c->Kown64);
    if (c->stress_stage) hipFree(c->stress_stage);
    free_two_level(*c);
    free_ilu0(*c);
    free_krylov_ws(*c);
    free_gnn_work(*c);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->asm_status) hipFree(c->asm_status);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace pdx

// ==========================================================================
// C-ABI
// ==========================================================================
using namespace pdx;

static thread_local std::string g_err;

namespace pdx {
int set_error(int rc, const std::string &msg) {
    g_err = msg;
    return rc;
}
}  // namespace pdx

static int fail(int rc, const std::string &msg) { return set_error(rc, msg); }

extern "C" {

static int explicit_driver(pdx_ctx *pc, int nsteps, int prec, int use_graph, std::string &err);

const char *pdx_last_error(void) { return g_err.c_str(); }
const char *pdx_version(void) { return "pdx 0.1 (gfx950)"; }

int pdx_set_setup_device(int device) {
    if (device < -1) return fail(PDX_ERR_ARG, "pdx_set_setup_device: device must be -1 or a GPU index");
    if (device >= 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || device >= n)
            return fail(PDX_ERR_ARG, "pdx_set_setup_device: no GPU " + std::to_string(device));
    }
    set_setup_device(device);
    return PDX_OK;
}

int pdx_model_create(const pdx_mesh_desc *desc, pdx_model **out) {
    if (!desc || !out) return fail(PDX_ERR_ARG, "null argument");
    auto *pm = new pdx_model();
    std::string err;
    int rc = model_from_desc(*desc, pm->m, err);
    if (rc) { delete pm; return fail(rc, err); }
    *out = pm;
    return PDX_OK;
}

int pdx_model_read(const char *cmd_file, pdx_model **out, int *solver_flag, int *num_steps) {
    if (!cmd_file || !out) return fail(PDX_ERR_ARG, "null argument");
    auto *pm = new pdx_model();
    std::string err;
    int sf = -1, ns = 1;
    int rc = model_read_files(cmd_file, pm->m, sf, ns, err);
    if (rc) { delete pm; return fail(rc, err); }
    if (solver_flag) *solver_flag = sf;
    if (num_steps) *num_steps = ns;
    *out = pm;
    return PDX_OK;
}

void pdx_model_destroy(pdx_model *m) { delete m; }

int pdx_write_vtk(const pdx_model *pm, const char *path, const char *title, const double *u, const double *sigma,
                  const double *phi, int binary) {
    if (!pm || !path) return fail(PDX_ERR_ARG, "pdx_write_vtk: null argument");
    std::string err;
    const int rc = write_vtk(pm->m, path, title ? std::string(title) : pm->m.title, u, sigma, phi, binary != 0, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_write_vtk_part(const pdx_model *pm, const char *path, const char *title, const double *u, const double *sigma,
                       const double *phi, int part) {
    if (!pm || !path) return fail(PDX_ERR_ARG, "pdx_write_vtk_part: null argument");
    std::string err;
    const int rc =
        write_vtk_part(pm->m, path, title ? std::string(title) : pm->m.title, u, sigma, phi, part, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_write_node_table(const pdx_model *pm, const char *path, const double *u, const double *sigma, int kind) {
    if (!pm || !path) return fail(PDX_ERR_ARG, "pdx_write_node_table: null argument");
    std::string err;
    const int rc = write_node_table(pm->m, path, u, sigma, kind, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_model_sizes(const pdx_model *pm, long long *out) {
    if (!pm || !out) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    int npresc = 0;
    for (int k = 0; k < m.nn; k++)
        for (int i = 0; i < m.dim; i++) npresc += m.eq[(size_t)k * 3 + i] < 0;
    long long v[16] = {m.nn, m.ne, m.dim, m.neq, (long long)m.nnz, m.nfam, (long long)m.fam_mem.size(), m.npdbe,
                       (long long)m.pde.size(), (long long)m.fe.size(), (long long)m.adj.size(), npresc,
                       m.nb[0], m.nb[1], m.nb[2], m.max_family};
    std::memcpy(out, v, sizeof(v));
    return PDX_OK;
}

int pdx_model_adjacency(const pdx_model *pm, int *adj_off, int *adj) {
    if (!pm) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    if (adj_off) std::copy(m.adj_off.begin(), m.adj_off.end(), adj_off);
    if (adj) std::copy(m.adj.begin(), m.adj.end(), adj);
    return PDX_OK;
}

int pdx_model_csr(const pdx_model *pm, long long *ia, long long *ja) {
    if (!pm) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    if (ia) std::copy(m.ia.begin(), m.ia.end(), ia);
    if (ja) {
        const std::vector<int64_t> &mj = m.csr_ja();
        std::copy(mj.begin(), mj.end(), ja);
    }
    return PDX_OK;
}

int pdx_model_families(const pdx_model *pm, int *fam_node, int *fam_off, int *fam_mem, int *bond, double *cd_v,
                       double *delta) {
    if (!pm) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    if (fam_node) std::copy(m.fam_node.begin(), m.fam_node.end(), fam_node);
    if (fam_off) std::copy(m.fam_off.begin(), m.fam_off.end(), fam_off);
    if (fam_mem)
        for (size_t i = 0; i < m.fam_mem.size(); i++) fam_mem[i] = m.fam_mem[i] + 1;
    if (bond) std::copy(m.bond.begin(), m.bond.end(), bond);
    if (cd_v) std::copy(m.fam_v.begin(), m.fam_v.end(), cd_v);
    if (delta) std::copy(m.fam_delta.begin(), m.fam_delta.end(), delta);
    return PDX_OK;
}

int pdx_model_nodes(const pdx_model *pm, int *eq, int *ntype, double *nv, double *ev, int *pdbe, int *algo) {
    if (!pm) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    if (eq) std::copy(m.eq.begin(), m.eq.end(), eq);
    if (ntype) std::copy(m.ntype.begin(), m.ntype.end(), ntype);
    if (nv) std::copy(m.nv.begin(), m.nv.end(), nv);
    if (ev) std::copy(m.ev.begin(), m.ev.end(), ev);
    if (pdbe)
        for (size_t i = 0; i < m.pdbe.size(); i++) pdbe[i] = m.pdbe[i] + 1;
    if (algo) std::copy(m.algo.begin(), m.algo.end(), algo);
    return PDX_OK;
}

int pdx_model_values(const pdx_model *pm, double *u_all) {
    if (!pm || !u_all) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    for (int k = 0; k < m.nn; k++)
        for (int i = 0; i < m.dim; i++) u_all[(size_t)k * m.dim + i] = m.dofval[(size_t)k * 3 + i];
    return PDX_OK;
}

int pdx_ctx_create(const pdx_model *pm, int device, pdx_ctx **out) {
    if (!pm || !out) return fail(PDX_ERR_ARG, "null argument");
    auto *pc = new pdx_ctx();
    std::string err;
    int rc = ctx_build(pm->m, device, pc->c, err);
    if (rc) {
        ctx_free(pc->c);
        delete pc;
        return fail(rc, err);
    }
    pc->model = pm;
    const Model &m = pm->m;
    pc->dofval.resize((size_t)m.nn * m.dim);
    pdx_model_values(pm, pc->dofval.data());
    pc->bond = m.bond;
    *out = pc;
    return PDX_OK;
}

void pdx_ctx_destroy(pdx_ctx *pc) {
    if (!pc) return;
    ctx_free(pc->c);
    delete pc;
}

void *pdx_ctx_stream(pdx_ctx *pc) { return pc && pc->c ? (void *)pc->c->stream : nullptr; }

int pdx_set_dof_values(pdx_ctx *pc, const double *u_all) {
    if (!pc || !u_all) return fail(PDX_ERR_ARG, "null argument");
    std::copy(u_all, u_all + pc->dofval.size(), pc->dofval.begin());
    return PDX_OK;
}

int pdx_set_dt(pdx_ctx *pc, double dt) {
    if (!pc || !(dt > 0)) return fail(PDX_ERR_ARG, "bad dt");
    pc->c->dt = dt;
    for (auto &g : pc->c->graph)
        if (g) { hipGraphExecDestroy(g); g = nullptr; }
    return PDX_OK;
}

// F = F_ext - K u (masked to free rows) for a full DOF vector u (host)
static int force_from(pdx_ctx *pc, const std::vector<double> &u_all, bool with_ext, double *F, std::string &err) {
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    const size_t n = (size_t)c.nn * c.dim;
    double *du, *dy;
    PDX_HIP(hipMalloc((void **)&du, n * 8));
    PDX_HIP(hipMalloc((void **)&dy, n * 8));
    PDX_HIP(hipMemcpyAsync(du, u_all.data(), n * 8, hipMemcpyHostToDevice, c.stream));
    if (dev_spmv(c, du, dy, PDX_FP64, true)) { err = "spmv failed"; return PDX_ERR_HIP; }
    std::vector<double> y(n), fu;
    if (with_ext && c.n_nbc) {
        fu.resize((size_t)c.n_nbc * n);
        PDX_HIP(hipMemcpyAsync(fu.data(), c.f_unit, fu.size() * 8, hipMemcpyDeviceToHost, c.stream));
    }
    PDX_HIP(hipMemcpyAsync(y.data(), dy, n * 8, hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(du);
    hipFree(dy);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < c.dim; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r < 0) continue;
            size_t o = (size_t)k * c.dim + i;
            double fe = 0.0;
            if (with_ext)
                for (int s = 0; s < c.n_nbc; s++) fe = fe + c.nbc_val[s] * fu[(size_t)s * n + o];
            F[r] = fe - y[o];
        }
    return PDX_OK;
}

int pdx_assemble(pdx_ctx *pc, int mode, const double *UN, double *Ku, double *F, int with_ext) {
    if (!pc || !F) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    std::string err;
    if (mode == PDX_MODE_NEWMARK && !UN) return fail(PDX_ERR_ARG, "Newmark assembly needs U_N");
    int rc = dev_assemble_K(c, err);
    if (rc) return fail(rc, err);
    const size_t n = (size_t)c.nn * c.dim;
    std::vector<double> u(n, 0.0);
    if (mode == PDX_MODE_STATIC) {
        for (int k = 0; k < c.nn; k++)
            for (int i = 0; i < c.dim; i++) {
                int ps = m.presc_set[(size_t)k * 3 + i];
                if (ps >= 0) u[(size_t)k * c.dim + i] = pc->dofval[(size_t)k * c.dim + i];
            }
    } else if (mode == PDX_MODE_NEWMARK) {
        std::copy(UN, UN + n, u.begin());
    } else {
        u = pc->dofval;
    }
    if ((rc = force_from(pc, u, with_ext != 0, F, err))) return fail(rc, err);
    if (Ku && mode != PDX_MODE_EXPLICIT) {
        double *dk;
        if (hipMalloc((void **)&dk, std::max<size_t>(c.nnz, 1) * 8) != hipSuccess) return fail(PDX_ERR_HIP, "alloc");
        dev_extract_csr(c, c.K, dk);
        hipMemcpyAsync(Ku, dk, c.nnz * 8, hipMemcpyDeviceToHost, c.stream);
        hipStreamSynchronize(c.stream);
        hipFree(dk);
    }
    return hipGetLastError() == hipSuccess ? PDX_OK : fail(PDX_ERR_HIP, "HIP error after assembly");
}

int pdx_assemble_timed(pdx_ctx *pc, double *ms) {
    if (!pc) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    std::string err;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return fail(PDX_ERR_HIP, "event");
    // the events bracket the kernels; the status check after them is not timed
    hipEventRecord(e0, c.stream);
    int rc = dev_assemble_K_launch(c, err);
    hipEventRecord(e1, c.stream);
    if (!rc) rc = dev_assemble_K_finish(c, err);
    hipEventSynchronize(e1);
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ms) *ms = t;
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_lumped_mass(pdx_ctx *pc, double *mout) {
    if (!pc || !mout) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    std::string err;
    int rc = dev_lumped_mass(c, err);
    if (rc) return fail(rc, err);
    std::vector<double> h((size_t)c.nn * c.dim);
    hipMemcpyAsync(h.data(), c.mlump, h.size() * 8, hipMemcpyDeviceToHost, c.stream);
    hipStreamSynchronize(c.stream);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < c.dim; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r >= 0) mout[r] = h[(size_t)k * c.dim + i];
        }
    return PDX_OK;
}

int pdx_explicit_run(pdx_ctx *pc, int nsteps, int precision, double *u_out, double *u_prev) {
    if (!pc || nsteps < 0) return fail(PDX_ERR_ARG, "bad argument");
    std::string err;
    int rc = explicit_driver(pc, nsteps, precision, 1, err);
    if (rc) return fail(rc, err);
    rc = dev_explicit_state(*pc->c, u_out, u_prev, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_explicit_steps(pdx_ctx *pc, int nsteps, int precision, int use_graph) {
    if (!pc || nsteps < 0) return fail(PDX_ERR_ARG, "bad argument");
    std::string err;
    int rc = explicit_driver(pc, nsteps, precision, use_graph, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_step_log(const pdx_ctx *pc, double *info, int *n) {
    if (!pc || !n) return fail(PDX_ERR_ARG, "null argument");
    *n = (int)(pc->step_log.size() / 3);
    if (info) std::copy(pc->step_log.begin(), pc->step_log.end(), info);
    return PDX_OK;
}

int pdx_model_settings(const pdx_model *pm, pdx_settings *out) {
    if (!pm || !out) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pm->m;
    pdx_settings s{};
    s.solver_flag = m.solver_flag;
    s.newmark = m.newmark;
    s.lumped_mass = m.lumped;
    s.num_steps = m.num_steps;
    s.save_freq = m.save_freq;
    s.vtk_binary = m.vtk_binary;
    s.fc = m.fc;
    s.topk = m.topk;
    s.n_rf = (int)m.rf_ids.size();
    for (int i = 0; i < s.n_rf && i < 16; i++) s.rf_ids[i] = m.rf_ids[i];
    s.dt = m.dt;
    s.beta = m.beta;
    s.gamma = m.gamma;
    s.horizon_factor = m.fac;
    s.nlf = m.nlf;
    s.sigma_ult = m.sig_ult;
    s.kic = m.kic;
    *out = s;
    return PDX_OK;
}

int pdx_model_mass_csr(const pdx_model *pm, long long *ia, long long *ja, long long *nnz) {
    if (!pm || !nnz) return fail(PDX_ERR_ARG, "null argument");
    std::vector<int64_t> hia, hja;
    mass_pattern(pm->m, hia, hja);
    *nnz = (long long)hja.size();
    if (ia) std::copy(hia.begin(), hia.end(), ia);
    if (ja) std::copy(hja.begin(), hja.end(), ja);
    return PDX_OK;
}

int pdx_mass_csr(pdx_ctx *pc, double *Mv) {
    if (!pc || !Mv) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    std::string err;
    hipSetDevice(c.device);
    int rc = dev_mass_on_K(c, err);
    if (rc) return fail(rc, err);
    std::vector<double> mk(std::max<int64_t>(m.nnz, 1));
    double *dk = nullptr;
    if (hipMalloc((void **)&dk, mk.size() * 8) != hipSuccess) return fail(PDX_ERR_HIP, "alloc");
    rc = dev_extract_csr(c, c.Mk, dk);
    if (!rc && hipMemcpyAsync(mk.data(), dk, m.nnz * 8, hipMemcpyDeviceToHost, c.stream) != hipSuccess) rc = PDX_ERR_HIP;
    if (hipStreamSynchronize(c.stream) != hipSuccess) rc = PDX_ERR_HIP;
    hipFree(dk);
    if (rc) return fail(rc, "HIP error in the mass assembly");
    // the mass pattern lies inside K's pattern (family self / FE neighbours of every node)
    std::vector<int64_t> hia, hja;
    mass_pattern(m, hia, hja);
    const std::vector<int64_t> &kja = m.csr_ja();
    for (int r = 0; r < m.neq; r++)
        for (int64_t q = hia[r]; q < hia[r + 1]; q++) {
            auto b = kja.begin() + m.ia[r], e = kja.begin() + m.ia[r + 1];
            auto it = std::lower_bound(b, e, hja[q]);
            if (it == e || *it != hja[q]) return fail(PDX_ERR_PATTERN, "mass entry outside the stiffness pattern");
            Mv[q] = mk[it - kja.begin()];
        }
    return PDX_OK;
}

int pdx_last_timing(pdx_ctx *pc, double *total_ms, double *per_step_ms) {
    if (!pc) return fail(PDX_ERR_ARG, "null argument");
    if (total_ms) *total_ms = pc->c->last_total_ms;
    if (per_step_ms) *per_step_ms = pc->c->last_step_ms;
    return PDX_OK;
}

int pdx_node_stresses(pdx_ctx *pc, const double *u_all, double *sigma) {
    if (!pc || !sigma) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    std::string err;
    const size_t n = (size_t)c.nn * c.dim;
    double *du = nullptr, *ds = nullptr;
    auto run = [&]() -> int {
        PDX_HIP(hipMalloc((void **)&ds, (size_t)c.nn * 6 * sizeof(double)));
        const double *src;
        if (u_all) {
            PDX_HIP(hipMalloc((void **)&du, n * sizeof(double)));
            PDX_HIP(hipMemcpyAsync(du, u_all, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
            src = du;
        } else {
            // the DOF values of the last stored step: free u_{n+1} with the prescribed values
            // that step's force used (the state buffer already holds the next step's)
            if (c.steps_done < 0) { err = "no explicit state yet: pass the DOF values"; return PDX_ERR_ARG; }
            int rc = dev_explicit_dofvals(c, &src, err);
            if (rc) return rc;
        }
        int rc = dev_node_stresses(c, src, ds, err);
        if (rc) return rc;
        PDX_HIP(hipMemcpyAsync(sigma, ds, (size_t)c.nn * 6 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        PDX_HIP(hipStreamSynchronize(c.stream));
        return PDX_OK;
    };
    int rc = run();
    if (du) hipFree(du);
    if (ds) hipFree(ds);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_node_stresses_noalloc(pdx_ctx *pc, const double *u_dev, double *sigma_dev) {
    if (!pc || !u_dev || !sigma_dev) return fail(PDX_ERR_ARG, "pdx_node_stresses_noalloc: null argument");
    Ctx &c = *pc->c;
    std::string err;
    if (hipSetDevice(c.device) != hipSuccess) return fail(PDX_ERR_HIP, "hipSetDevice failed");
    const int rc = dev_node_stresses_noalloc(c, u_dev, sigma_dev, err);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_node_stresses_timed(pdx_ctx *pc, const double *u_all, double *ms) {
    if (!pc || !u_all || !ms) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    std::string err;
    const size_t n = (size_t)c.nn * c.dim;
    double *du = nullptr, *ds = nullptr;
    auto run = [&]() -> int {
        PDX_HIP(hipMalloc((void **)&ds, (size_t)c.nn * 6 * sizeof(double)));
        PDX_HIP(hipMalloc((void **)&du, n * sizeof(double)));
        PDX_HIP(hipMemcpyAsync(du, u_all, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
        int rc = dev_node_stresses(c, du, ds, err);  // warm (assembles K if needed)
        if (rc) return rc;
        hipEvent_t e0, e1;
        PDX_HIP(hipEventCreate(&e0));
        PDX_HIP(hipEventCreate(&e1));
        PDX_HIP(hipEventRecord(e0, c.stream));
        rc = dev_node_stresses(c, du, ds, err, false);
        PDX_HIP(hipEventRecord(e1, c.stream));
        PDX_HIP(hipEventSynchronize(e1));
        if (!rc) rc = stress_status_check(c, err);
        float t = 0;
        PDX_HIP(hipEventElapsedTime(&t, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        *ms = t;
        return rc;
    };
    int rc = run();
    if (du) hipFree(du);
    if (ds) hipFree(ds);
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_explicit_tuning(pdx_ctx *pc, int precision, int *lanes, int *nontemporal, int *xcd_order, int *row_soa,
                        int *col_classes, int *mirrored, int *alternating, int *form, float *ms) {
    if (!pc || (precision != PDX_FP64 && precision != PDX_FP32)) return fail(PDX_ERR_ARG, "bad argument");
    const Ctx &c = *pc->c;
    const int v = c.tune_explicit[precision];
    const ExplicitForm f = v >= 0 ? kExplicitForms[v] : ExplicitForm{0, false, false, false, false, false, false, 0};
    if (lanes) *lanes = f.lanes;
    if (nontemporal) *nontemporal = f.nt;
    if (xcd_order) *xcd_order = f.xm;
    if (row_soa) *row_soa = f.soa;
    if (col_classes) *col_classes = f.rel;
    if (mirrored) *mirrored = f.sym;
    if (alternating) *alternating = f.alt;
    if (form) *form = v;
    if (ms) std::copy(c.tune_ms[precision], c.tune_ms[precision] + PDX_EXPLICIT_FORMS, ms);
    return PDX_OK;
}

int pdx_explicit_set_form(pdx_ctx *pc, int precision, int form) {
    if (!pc || (precision != PDX_FP64 && precision != PDX_FP32) || form < -1 || form >= PDX_EXPLICIT_FORMS)
        return fail(PDX_ERR_ARG, "pdx_explicit_set_form: bad argument");
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    c.tune_explicit[precision] = form;
    if (c.graph[precision]) {  // the captured steps launch the previous form
        hipStreamSynchronize(c.stream);
        hipGraphExecDestroy(c.graph[precision]);
        c.graph[precision] = nullptr;
    }
    return PDX_OK;
}

int pdx_max_eig(pdx_ctx *pc, int iters, double *lmax) {
    if (!pc || !lmax || iters <= 0) return fail(PDX_ERR_ARG, "bad argument");
    std::string err;
    int rc = dev_max_eig(*pc->c, iters, lmax, err);
    return rc ? fail(rc, err) : PDX_OK;
}

static pdx_solve_opts default_opts() {
    pdx_solve_opts o{};
    o.krylov = PDX_KRYLOV_BICGSTAB;
    o.precond = PDX_PREC_JACOBI;
    o.precision = PDX_FP64;
    o.tol = 1e-8;
    o.max_iter = 20000;
    o.inner_tol = 1e-4;
    o.warm_start = 0;
    return o;
}

int pdx_static_solve(pdx_ctx *pc, const pdx_solve_opts *op, double *u, pdx_solve_stats *st) {
    if (!pc || !u) return fail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    pdx_solve_opts o = op ? *op : default_opts();
    std::string err;
    int rc;
    if (!c.K_valid && (rc = dev_assemble_K(c, err))) return fail(rc, err);
    const size_t n = (size_t)c.nn * c.dim;
    // b = F_ext - K_fp u_p, as a full DOF vector with prescribed entries zero
    std::vector<double> up(n, 0.0), Fq(std::max(c.neq, 1));
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < c.dim; i++)
            if (m.presc_set[(size_t)k * 3 + i] >= 0) up[(size_t)k * c.dim + i] = pc->dofval[(size_t)k * c.dim + i];
    if ((rc = force_from(pc, up, true, Fq.data(), err))) return fail(rc, err);
    std::vector<double> b(n, 0.0), x(n, 0.0);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < c.dim; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r >= 0) {
                b[(size_t)k * c.dim + i] = Fq[r];
                if (o.warm_start) x[(size_t)k * c.dim + i] = u[r];
            }
        }
    double *db, *dx;
    if (hipMalloc((void **)&db, n * 8) != hipSuccess || hipMalloc((void **)&dx, n * 8) != hipSuccess)
        return fail(PDX_ERR_HIP, "alloc");
    hipMemcpyAsync(db, b.data(), n * 8, hipMemcpyHostToDevice, c.stream);
    hipMemcpyAsync(dx, x.data(), n * 8, hipMemcpyHostToDevice, c.stream);
    pdx_solve_stats s{};
    rc = dev_solve(c, o, db, dx, s, err);
    hipMemcpyAsync(x.data(), dx, n * 8, hipMemcpyDeviceToHost, c.stream);
    hipStreamSynchronize(c.stream);
    hipFree(db);
    hipFree(dx);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < c.dim; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r >= 0) u[r] = x[(size_t)k * c.dim + i];
        }
    if (st) *st = s;
    return rc ? fail(rc, err) : PDX_OK;
}

// ---- quasi-static load steps with bond failure ----------------------------

// One step of failureCriterion_maxPriSig + Finally_TopK_andUpdate_bondStaus (flag 2)
// (pdsolve.cpp:7184-7289, :6981-7182) on the bond states of the context: the largest principal
// stress s1 of every intact bond (device), the bonds with s1 > sigma_ult kept as the topk largest
// (0: all; ties keep the bond met first in family / member order, as the reference's min-heap
// replaces only on a larger value), and for each kept bond a crack segment of length 2 * minD
// (minD = half the smallest horizon) through the bond's midpoint, normal to the principal
// direction, which breaks every intact bond that crosses it.  Returns the bonds kept; smax gets
// the largest s1 over the intact bonds (at least 0), nbroke the bonds broken.
static int fc2_step(pdx_ctx *pc, const double *sig_dev, const std::vector<double> &sig, double sigma_ult, int topk,
                    double &smax, int &nkept, int &nbroke, std::string &err) {
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    const size_t nb = pc->bond.size();
    double *s1_dev;
    int *bond_dev = c.bond;
    PDX_HIP(hipMalloc((void **)&s1_dev, std::max<size_t>(nb, 1) * 8));
    int rc = dev_bond_sigma1(c, bond_dev, sig_dev, s1_dev, err);
    std::vector<double> s1(nb);
    if (!rc && nb) {
        PDX_HIP(hipMemcpyAsync(s1.data(), s1_dev, nb * 8, hipMemcpyDeviceToHost, c.stream));
        PDX_HIP(hipStreamSynchronize(c.stream));
    }
    hipFree(s1_dev);
    if (rc) return rc;
    smax = 0.0;
    struct Cand {
        double v;
        long long order;  // family / member scan order
        int f, q;
    };
    std::vector<Cand> cand;
    long long order = 0;
    for (int f = 0; f < m.nfam; f++)
        for (int q = m.fam_off[f] + 1; q < m.fam_off[f + 1]; q++, order++) {
            if (pc->bond[q] != 1) continue;
            const double v = s1[q];
            if (v > smax) smax = v;
            if (v > sigma_ult) cand.push_back({v, order, f, q});
        }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.v > b.v; });
    if (topk > 0 && (int)cand.size() > topk) cand.resize(topk);
    nkept = (int)cand.size();
    nbroke = 0;
    const double minD = 0.5 * m.minDelta;
    const std::vector<int> bond_before = cand.empty() ? std::vector<int>() : pc->bond;
    for (const Cand &k : cand) {
        const int nk = m.fam_node[k.f], nm = m.fam_mem[k.q];
        const double *xk = &m.x[(size_t)nk * 3], *xm = &m.x[(size_t)nm * 3];
        double xc[3];
        for (int i = 0; i < 3; i++) xc[i] = 0.5 * (xk[i] + xm[i]);
        const double *sk = &sig[(size_t)nk * 6], *sm = &sig[(size_t)nm * 6];
        const double a = 0.5 * (sk[0] + sm[0]), b = 0.5 * (sk[1] + sm[1]), cc = 0.5 * (sk[3] + sm[3]);
        // unit eigenvector of the larger eigenvalue of [[a, cc], [cc, b]] (the first axis for a
        // multiple of the identity)
        const double h = 0.5 * (a - b), l1 = 0.5 * (a + b) + std::sqrt(h * h + cc * cc);
        double e0, e1;
        if (cc != 0.0) {
            e0 = l1 - b;
            e1 = cc;
        } else if (a >= b) {
            e0 = 1.0;
            e1 = 0.0;
        } else {
            e0 = 0.0;
            e1 = 1.0;
        }
        const double nrm = std::sqrt(e0 * e0 + e1 * e1);
        e0 /= nrm;
        e1 /= nrm;
        double crack[9] = {xc[0] - minD * e1, xc[1] + minD * e0, 0.0, xc[0] + minD * e1, xc[1] - minD * e0, 0.0,
                           0.0, 0.0, 0.0};
        nbroke += break_bonds_across(m, crack, pc->bond);
    }
    if (nbroke) {
        PDX_HIP(hipMemcpy(c.bond, pc->bond.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        // the next assembly sees the new bond states: the families with a changed bond get their
        // rows of K recomputed
        for (int f = 0; f < m.nfam; f++)
            for (int q = m.fam_off[f] + 1; q < m.fam_off[f + 1]; q++)
                if (pc->bond[q] != bond_before[q]) {
                    c.dirty_fams.push_back(f);
                    break;
                }
        c.K_valid = false;
    }
    return PDX_OK;
}

// calGlobalNodeStresses + failureProcess after a stored step of a dynamic driver
// (pdsolve.cpp:6060-6064 explicit, :6258-6261 Newmark): nodal stresses of the DOF values, then
// the FC 2 update (flag 2, non-static solver: the TopK bonds above the strength break, with the
// crack segments of Finally_TopK_andUpdate_bondStaus); broken bonds invalidate K, which the
// driver assembles again before the next step's force.  Logs (kept, broken, largest s1).
static int dynamic_failure(pdx_ctx *pc, const double *uall_dev, std::string &err) {
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    double *dsig = nullptr;
    PDX_HIP(hipMalloc((void **)&dsig, (size_t)c.nn * 6 * 8));
    std::vector<double> sig((size_t)c.nn * 6);
    int rc = dev_node_stresses(c, uall_dev, dsig, err);
    if (!rc && hipMemcpyAsync(sig.data(), dsig, sig.size() * 8, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        rc = PDX_ERR_HIP;
    if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) rc = PDX_ERR_HIP;
    double smax = 0.0;
    int nkept = 0, nbroke = 0;
    if (!rc) rc = fc2_step(pc, dsig, sig, m.sig_ult, m.topk, smax, nkept, nbroke, err);
    hipFree(dsig);
    if (rc) return rc;
    pc->step_log.push_back(nkept);
    pc->step_log.push_back(nbroke);
    pc->step_log.push_back(smax);
    return PDX_OK;
}

// failure criteria of the dynamic drivers: FC 0 none, FC 2 in 2D (failureProcess returns at once
// in 3D, pdsolve.cpp:6594-6597); FC 1 (the J-integral crack growth) is not built
static int dynamic_failure_mode(const Model &m, bool &fail2, std::string &err) {
    fail2 = false;
    if (m.fc != 0 && m.fc != 2) {
        err = "failure criterion FC " + std::to_string(m.fc) +
              " is not supported by the dynamic drivers (FC 0 and FC 2 are)";
        return PDX_ERR_ARG;
    }
    fail2 = m.fc == 2 && m.dim == 2;
    if (fail2 && !(m.sig_ult > 0)) {
        err = "FC 2 needs the material strength (sigma_ult > 0)";
        return PDX_ERR_ARG;
    }
    return PDX_OK;
}

// pdfemDynamicSolver_CSRformat (pdsolve.cpp:5842-6090): lumped or consistent mass as the model's
// LUMPMASS says, FC 2 bond failure after every stored step
static int explicit_driver(pdx_ctx *pc, int nsteps, int prec, int use_graph, std::string &err) {
    if (prec != PDX_FP64 && prec != PDX_FP32) { err = "precision must be PDX_FP64 or PDX_FP32"; return PDX_ERR_ARG; }
    const Model &m = pc->model->m;
    bool fail2;
    if (int rc = dynamic_failure_mode(m, fail2, err)) return rc;
    hipSetDevice(pc->c->device);
    std::function<int(std::string &)> cb;
    if (fail2)
        cb = [pc](std::string &e) -> int {
            const double *u = nullptr;
            int rc = dev_explicit_dofvals(*pc->c, &u, e);
            return rc ? rc : dynamic_failure(pc, u, e);
        };
    return dev_explicit_steps(*pc->c, nsteps, prec, use_graph, m.lumped, cb, err);
}

int pdx_newmark_run(pdx_ctx *pc, int nsteps, const pdx_solve_opts *op, double *u, double *v, double *a,
                    pdx_solve_stats *st) {
    if (!pc || nsteps < 0) return fail(PDX_ERR_ARG, "bad argument");
    pdx_solve_opts o = op ? *op : default_opts();
    if (!op) o.warm_start = 1;
    std::string err;
    bool fail2;
    if (int rc = dynamic_failure_mode(pc->model->m, fail2, err)) return fail(rc, err);
    hipSetDevice(pc->c->device);
    std::function<int(const double *, std::string &)> cb;
    if (fail2) cb = [pc](const double *uall, std::string &e) -> int { return dynamic_failure(pc, uall, e); };
    pdx_solve_stats s{};
    int rc = dev_newmark(*pc->c, nsteps, o, pc->dofval, u, v, a, s, cb, err);
    if (st) *st = s;
    return rc ? fail(rc, err) : PDX_OK;
}

int pdx_quasi_static_run(pdx_ctx *pc, int nsteps, const pdx_solve_opts *op, int fc, int topk, double sigma_ult,
                         double *u_all, double *step_info, pdx_solve_stats *st) {
    if (!pc || nsteps < 0) return fail(PDX_ERR_ARG, "pdx_quasi_static_run: bad argument");
    Ctx &c = *pc->c;
    const Model &m = pc->model->m;
    const pdx_solve_opts o = op ? *op : default_opts();
    if (fc < 0) fc = m.fc;
    if (topk < 0) topk = m.topk;
    if (!(sigma_ult > 0)) sigma_ult = m.sig_ult;
    if (fc != 0 && fc != 2) return fail(PDX_ERR_ARG, "pdx_quasi_static_run: failure criterion 0 or 2");
    if (fc == 2 && m.dim == 2 && !(sigma_ult > 0)) return fail(PDX_ERR_ARG, "pdx_quasi_static_run: sigma_ult <= 0");
    std::string err;
    int rc = 0;
    const size_t n = (size_t)c.nn * c.dim;
    const double dt = m.dt;
    std::vector<double> ebc_cur(m.ebc.size());
    for (size_t s = 0; s < m.ebc.size(); s++) ebc_cur[s] = m.ebc[s].val;
    // current prescribed values are the DOF values of the context
    for (size_t s = 0; s < m.ebc.size(); s++)
        if (!m.ebc[s].nid.empty()) ebc_cur[s] = pc->dofval[(size_t)m.ebc[s].nid[0] * c.dim + m.ebc[s].dof];
    double *db = nullptr, *dx = nullptr, *du = nullptr, *dsig = nullptr;
    auto cleanup = [&]() {
        for (double *p : {db, dx, du, dsig})
            if (p) hipFree(p);
    };
    if (hipMalloc((void **)&db, n * 8) != hipSuccess || hipMalloc((void **)&dx, n * 8) != hipSuccess ||
        hipMalloc((void **)&du, n * 8) != hipSuccess || hipMalloc((void **)&dsig, (size_t)c.nn * 6 * 8) != hipSuccess) {
        cleanup();
        return fail(PDX_ERR_HIP, "alloc");
    }
    std::vector<double> Fq(std::max(c.neq, 1)), b(n), x(n), sig((size_t)c.nn * 6);
    pdx_solve_stats tot{};
    bool add_load = true;
    for (int step = 0; step < nsteps && !rc; step++) {
        // pdfemQuasiStaticAssembleEquasSys_CSRformat (pdsolve.cpp:6434-6462): K_ff and
        // F = F_ext - K u over all columns at the current DOF values; K_ff du = F
        if (!c.K_valid && (rc = dev_assemble_K(c, err))) break;
        if ((rc = force_from(pc, pc->dofval, true, Fq.data(), err))) break;
        std::fill(b.begin(), b.end(), 0.0);
        std::fill(x.begin(), x.end(), 0.0);
        for (int k = 0; k < c.nn; k++)
            for (int i = 0; i < c.dim; i++) {
                const int r = m.eq[(size_t)k * 3 + i];
                if (r >= 0) b[(size_t)k * c.dim + i] = Fq[r];
            }
        bool loaded = false;
        for (double v : b) loaded = loaded || v != 0.0;
        if (loaded) {  // an unloaded step (F = 0) has du = 0
            hipMemcpyAsync(db, b.data(), n * 8, hipMemcpyHostToDevice, c.stream);
            hipMemcpyAsync(dx, x.data(), n * 8, hipMemcpyHostToDevice, c.stream);
            pdx_solve_stats s{};
            if ((rc = dev_solve(c, o, db, dx, s, err))) break;
            tot.iterations += s.iterations;
            tot.refinements += s.refinements;
            tot.true_relres = s.true_relres;
            tot.seconds += s.seconds;
            tot.converged = s.converged;
            hipMemcpyAsync(x.data(), dx, n * 8, hipMemcpyDeviceToHost, c.stream);
            hipStreamSynchronize(c.stream);
        }
        for (int k = 0; k < c.nn; k++)
            for (int i = 0; i < c.dim; i++)
                if (m.eq[(size_t)k * 3 + i] >= 0) pc->dofval[(size_t)k * c.dim + i] += x[(size_t)k * c.dim + i];
        // calGlobalNodeStresses, then failureProcess (pdsolve.cpp:6601-6627; 3D returns at once)
        double rati = 1.0, smax = 0.0;
        int nkept = 0, nbroke = 0;
        if (fc == 2 && c.dim == 2) {
            hipMemcpyAsync(du, pc->dofval.data(), n * 8, hipMemcpyHostToDevice, c.stream);
            if ((rc = dev_node_stresses(c, du, dsig, err))) break;
            hipMemcpyAsync(sig.data(), dsig, sig.size() * 8, hipMemcpyDeviceToHost, c.stream);
            hipStreamSynchronize(c.stream);
            if ((rc = fc2_step(pc, dsig, sig, sigma_ult, topk, smax, nkept, nbroke, err))) break;
            add_load = nkept == 0;
            rati = std::fabs(smax) < 1e-20 * sigma_ult ? 1.0 : 1.001 * sigma_ult / smax;
        }
        if (step_info) {
            step_info[step * 4] = nkept;
            step_info[step * 4 + 1] = nbroke;
            step_info[step * 4 + 2] = rati;
            step_info[step * 4 + 3] = smax;
        }
        if (add_load) {
            // setVaryNaturalBC, setVaryEssentialBC with the ratio (pdsolve.cpp:5263-5329)
            for (int s2 = 0; s2 < c.n_nbc; s2++)
                if (c.nbc_vary[s2]) c.nbc_val[s2] = c.nbc_val[s2] + dt * c.nbc_vel[s2];
            for (size_t s2 = 0; s2 < m.ebc.size(); s2++) {
                const EbcSet &E = m.ebc[s2];
                if (!E.vary) continue;
                double dv;
                if (ebc_cur[s2] == 0) {
                    dv = E.vel;
                } else {
                    const double ndv = (rati - 1) * ebc_cur[s2];
                    if (std::fabs(ndv) > std::fabs(E.vel)) dv = E.vel;
                    else if (rati < 1) dv = 0;
                    else dv = ndv;
                }
                ebc_cur[s2] = ebc_cur[s2] + dt * dv;
                for (int k : E.nid) pc->dofval[(size_t)k * c.dim + E.dof] = ebc_cur[s2];
            }
        }
    }
    cleanup();
    if (rc) return fail(rc, err);
    if (u_all) std::copy(pc->dofval.begin(), pc->dofval.end(), u_all);
    if (st) *st = tot;
    return PDX_OK;
}

int pdx_ctx_bonds(pdx_ctx *pc, int *bond) {
    if (!pc || !bond) return fail(PDX_ERR_ARG, "null argument");
    std::copy(pc->bond.begin(), pc->bond.end(), bond);
    return PDX_OK;
}

int pdx_set_incremental_bonds(pdx_ctx *pc, int on) {
    if (!pc) return fail(PDX_ERR_ARG, "null argument");
    pc->c->incremental_bonds = on != 0;
    return PDX_OK;
}

int pdx_local_damage(pdx_ctx *pc, double *phi) {
    // calLocalDamage (pdsolve.cpp:7338-7373): 1 - intact / all bonds of the node's family
    if (!pc || !phi) return fail(PDX_ERR_ARG, "null argument");
    const Model &m = pc->model->m;
    std::fill(phi, phi + m.nn, 0.0);
    for (int f = 0; f < m.nfam; f++) {
        double t = 0;
        for (int q = m.fam_off[f] + 1; q < m.fam_off[f + 1]; q++) t = t + pc->bond[q];
        phi[m.fam_node[f]] = 1.0 - t / (m.fam_off[f + 1] - m.fam_off[f] - 1.0);
    }
    return PDX_OK;
}

}  // extern "C"
