// Library context: problem data, sector dimensions, the device operator pool and the
// per-site Heff work lists.
#pragma once
#include <thread>
#include <tuple>

#include "chm_internal.hpp"
#include "dbuf.hpp"
#include "heff_plan.hpp"
#include "learned.hpp"
#include "update_plan.hpp"

namespace chm {

enum { KIND_L = 0, KIND_F0, KIND_F1, KIND_S0, KIND_S1, KIND_A, KIND_B, KIND_C, KIND_D, KIND_Q, KIND_X, KIND_COUNT };

struct OpShape {
  int kind, k, m, two_j, n_elec, n_irrep;
};

// Page-locked host staging memory that only grows (hipHostMalloc); sub-allocated in
// order by take() after reset().
struct HostStage {
  char* p = nullptr;
  size_t n = 0, used = 0;
  HostStage() = default;
  HostStage(const HostStage&) = delete;
  HostStage& operator=(const HostStage&) = delete;
  ~HostStage() { if (p) (void)hipHostFree(p); }
  void reserve(size_t bytes) {
    if (bytes <= n) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    const size_t want = bytes + bytes / 4;
    hip_check(hipHostMalloc((void**)&p, want, hipHostMallocDefault), "hipHostMalloc staging");
    n = want;
  }
  void reset() { used = 0; }
  template <class T>
  T* take(size_t count) {
    const size_t off = (used + 255) / 256 * 256;
    if (off + count * sizeof(T) > n) throw Error{ERR_STATE, "staging buffer too small"};
    used = off + count * sizeof(T);
    return (T*)(p + off);
  }
};

// hipMemcpy after the context's stream has drained: the stream does not synchronise
// with the null stream, so a plain hipMemcpy could overtake kernels still using the
// destination or the source.
inline void copy_sync(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind, const char* what) {
  hip_check(hipStreamSynchronize(s), what);
  hip_check(hipMemcpy(dst, src, bytes, kind), what);
}

// Host-side state: the problem, the sector dimensions and the operator layouts.
struct HostState {
  // problem (DMRG orbital order)
  int L = 0, group = 0, nirrep = 1, N = 0, TwoS = 0, Irrep = 0;
  std::vector<int> irreps;
  std::vector<double> mx;
  double econst = 0.0;
  bool have_problem = false;
  Sectors sec;
  std::vector<OpSet> sets;  // slot cnt -> operator set
  std::vector<std::vector<Sectors::Sec>> fci;  // [b] FCI dimensions (> 0), when known
  // Throws unless slot cnt holds every operator DMRG::allocateTensors creates there.
  void validate_set(int cnt, bool mr) const;
};

// Overlap operator between the current MPS (up) and one lower state (down) at one
// boundary: TensorO (TensorO.cpp:27-193), the identity operator (two_j = n_elec =
// n_irrep = 0) with up dimensions from the current sectors and down dimensions from the
// lower state's; one block per sector where both are non-zero, in the order of the
// current sector list.
struct OvlT {
  int boundary = -1, dir = 0;  // dir 1 built moving right, 2 moving left
  int32_t buf = -1;
  int64_t size = 0;
  std::vector<int> pos2blk;    // position in the current list of the boundary -> block or -1
  std::vector<int> N, S, I, du, dd;
  std::vector<int64_t> k2i;
  Ref blk(const Sectors& cur, int n, int s, int i) const {
    const int p = cur.idx[boundary]->find(n, s, i);
    if (p < 0 || p >= (int)pos2blk.size() || pos2blk[p] < 0) return Ref{};
    return Ref{buf, k2i[pos2blk[p]]};
  }
};

// A converged lower state kept for excited-state DMRG (DMRG::newExcitation,
// DMRG.cpp:475-505): its energy shift, virtual dimensions, MPS on the device and its
// overlap operators with the current MPS per slot (Exc_Eshifts, Exc_BKs, Exc_MPSs,
// Exc_Overlaps).
struct LowerState {
  double eshift = 0.0;
  Sectors sec;
  std::vector<TLay> mps;
  std::vector<OvlT> ovl;  // [slot cnt] -> boundary cnt + 1
};

struct Ctx : HostState {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // device buffers by id (ids >= BUF_TENSOR0 are operator tensors)
  std::vector<double*> bufs;
  std::vector<int64_t> buf_len;

  // prepared site
  int site = -1;
  SLayout S;
  std::unique_ptr<HeffPlanner> plan;
  // Work lists of the next site generated on a background thread while the GPU runs the
  // current one (used while the sector dimensions stay fixed); valid only if the sectors
  // are unchanged when prepare() takes them.
  struct PlanJob {
    int site = -1;
    uint64_t gen = 0;
    SLayout S;
    std::unique_ptr<HeffPlanner> plan;
    std::thread th;
    std::exception_ptr err;
  };
  std::unique_ptr<PlanJob> job;
  void start_plan(int site);
  void drop_plan();
  std::vector<char> prealloc;     // [slot]: laid out by alloc_slot for the coming update
  void alloc_slot(int index, bool moving_right);
  // Work lists are released on a background thread while the GPU runs the step (tens
  // of millions of records); the previous release is joined first.
  std::thread reaper;
  template <class T>
  void reap(std::unique_ptr<T> p) {
    if (reaper.joinable()) reaper.join();
    reaper = std::thread([q = std::move(p)]() mutable { q.reset(); });
  }
  void reap_join() {
    if (reaper.joinable()) reaper.join();
  }
  double F_H = 0.0, B_H = 0.0;
  int64_t V = 0;
  DBuf<double> combo_arena, W, dx, dy, partial, scal;
  DBuf<DevTile1> tiles1;  // phase 1: W tiles of the two-operand terms, each with its product
  DBuf<DevTile> tiles2;
  std::vector<int> tb1, tb2;  // tile ranges of the batches that share the W scratch
  DBuf<DevBRed> treds;        // phase-2 output blocks split into segments: their reductions
  int ntreds = 0;
  int64_t tpart_off = 0;      // first partial sum of those segments in the chained partials
  int tile_segments = 8192;   // CHM_TILE_SEGMENTS: phase-2 workgroups aimed at per launch
  int tile_seg_min = 64;      // CHM_TILE_SEG_MIN: fewest 16-deep k chunks per segment
  bool wave_tiles = false;    // CHM_WAVE_TILES=1: every tile 32 x 32 on one wavefront (no LDS sharing)
  int64_t tile_w_cap = 0;  // doubles of W scratch per batch (CHM_TILE_W_CAP; 0: a quarter of free memory)
  DBuf<DevGemmOp> ops2;
  int ntiles1 = 0, ntiles2 = 0;
  DBuf<DevDiagBlock> dblocks;
  std::vector<DevDiagBlock> hblocks;  // block outputs as offsets into the diagonal
  DBuf<DevDiag> dterms;
  DBuf<int64_t> d_k2i;
  DBuf<double> d_sym, d_prog;  // per-block scale factors
  int64_t stats[16] = {0};
  // explicit block-sparse part of Heff (pairs with |kappa||kappa'| <= threshold)
  int64_t explicit_max = 4096;       // CHM_EXPLICIT_MAX
  double explicit_budget_gb = 0.0;   // CHM_EXPLICIT_BUDGET_GB (0: 40 % of free memory)
  int64_t explicit_threshold = 0;
  DBuf<double> Mvals;
  DBuf<int> Mcol;
  // Heff forms: 0 FP64; 1 mixed: the explicit matrix stored as hi/lo float pairs and the
  // fused chained items in FP32 (k_fchain on FP32 MFMA); the Davidson iterates with these
  // low-precision products (hi floats only) until the residual falls below mixed_switch
  // or stops falling, then restarts from the Ritz vector with FP64 products (hi + lo,
  // 2^-48) to the tolerance.
  int heff_form = 0;
  bool M_split = false;              // storage of Mvals at the prepared site
  int64_t M_nnz = 0;
  bool mv_low = false;               // apply() reads the hi floats only
  // Lower states' projector vectors (VeffTilde, already scaled by sqrt(E_shift)) for the
  // prepared site: apply() adds sum_q (v_q . x) v_q, diag() adds sum_q v_q^2. Cleared by
  // prepare().
  int n_lower = 0;
  DBuf<double> lower;                // n_lower x V
  DBuf<const double*> lower_ptrs;
  DBuf<double> lower_part, lower_dot;
  double mixed_switch = 1e-5;
  // Excited states: the lower states, the buffer id of the lower-state vectors and the
  // lower two-site object joined at each site.
  std::vector<LowerState> lowers;
  int32_t low_buf = -1, sup_buf = -1;
  // device buffer id with at least n doubles (allocated or grown; owned by bufs)
  double* own_buf(int32_t& id, int64_t n);
  void new_excitation(double eshift, int D, uint64_t seed);
  void update_overlaps(int index, bool mr);
  void veff_tilde(int index);
  bool ovl_current(const OvlT& o, int dir) const;
  void drop_lowers();
  // 2-RDM (twodm.cpp): the P tensors of the site being filled, the dot-product records
  // and their results, and the seconds of the last 2-RDM sweep
  int32_t rdm_pbuf = -1;
  DBuf<DevDot> rdm_items;
  DBuf<double> rdm_res;
  double rdm_seconds = 0.0;
  void rdm_fill_site(int site, std::vector<double>& A, std::vector<double>& B);
  double two_rdm(double* A, double* B);
  int64_t n_low_mults = 0, n_full_mults = 0;
  int64_t n_learned_fallbacks = 0;  // Davidson solves where the learned step was switched off
  // Learned preconditioner (learned.hip): precond_mode 1 multiplies the diagonal
  // preconditioner by (1 + m) with m from the MLP; harvest_per_iter > 0 records, at
  // every preconditioner application, the features, the plain preconditioned residual p
  // and (at convergence) the ideal correction e of that many elements.
  Mlp mlp;
  int precond_mode = 0;
  int harvest_per_iter = 0;
  DBuf<float> eldims, featX, hv_rows;
  DBuf<int64_t> hv_idx;
  int eldims_site = -1;
  uint64_t eldims_gen = 0;
  std::vector<float> hv_feats, hv_p, hv_e, hv_w;
  uint64_t hv_counter = 0;
  void build_eldims();
  // Autotuned routing (chm_set_tuning): rows keyed by the site's largest block dimension
  // and V select explicit_max, chain_item_terms, chain_group_terms, use_chain_mfma and
  // fchain_min for that site; sites no row covers use the defaults (chm_set_heff_knobs, the
  // CHM_* environment knobs or the built-in values).
  struct TuneRow {
    int64_t dmax_lo, dmax_hi, v_lo, v_hi, explicit_max, item_terms, group_terms, chain_mfma, fchain_min;
  };
  // knobs the environment set at creation (explicit_max, item_terms, group_terms,
  // chain_mfma, fchain_min): the table does not override them
  bool env_pinned[5] = {false, false, false, false, false};
  std::vector<TuneRow> tune;
  int64_t def_explicit_max = 4096;
  int def_item_terms = CHAIN_ITEM_TERMS, def_group_terms = CHAIN_GROUP_TERMS;
  bool def_chain_mfma = false;
  int tune_row = -1;  // row used at the last prepare (-1: defaults)
  int64_t site_dmax = 0;
  void apply_tuning();
  int use_chain_kernel = -1;  // CHM_CHAIN_KERNEL: 1 chained kernels for blocks up to 64, 0 GEMM tiles only, -1 by form
  // CHM_CHAIN_MFMA=1: chained blocks up to 48 on the matrix cores (k_chain_mfma). Off by
  // default: at L = 32, D = 2000 it matches k_chain for blocks of 33-48 and is slower
  // below (its operand loads from global memory are not coalesced).
  bool use_chain_mfma = false;
  DBuf<DevCBlock> cblocks_m;         // k_chain_mfma items sorted by tile count
  bool svd_large_gesvdj = true;      // CHM_SVD_GESVDJ=0: large Split sectors by rocsolver_dgesvd instead of dgesvdj
  int chain_item_terms = CHAIN_ITEM_TERMS;    // CHM_CHAIN_ITEM_TERMS
  int chain_group_terms = CHAIN_GROUP_TERMS;  // CHM_CHAIN_GROUP_TERMS
  int tiles_kernel = 2;  // CHM_TILES_MFMA: grouped GEMM tiles (Heff phase 2, updates) in 2 k_tiles_fit, 1 k_tiles_mfma, 0 k_tiles
  int cblocks_m_offs[CHAIN_MFMA_NT + 1] = {};
  DBuf<DevCBlock> cblocks;
  DBuf<DevCGroup> cgroups;
  DBuf<DevCTerm> cterms;
  int ncblocks = 0;
  int chain_P = 1;  // doubles per LDS array of the chained kernel
  // k_fchain (chained products of blocks up to FCHAIN_MAXD on the matrix cores): work
  // items in buckets of tile shape (fchain_bucket). CHM_FCHAIN=0
  // turns it off; blocks whose largest dimension is at most fchain_min (CHM_FCHAIN_MIN)
  // stay on k_chain.
  bool use_fchain = true, fchain_ok = false;
  // the mixed form's low-precision products run k_fchain's items in FP32 on the matrix
  // cores; CHM_FCHAIN_FP32=0 keeps them FP64
  bool fchain_fp32 = true;
  int fchain_min = 32, def_fchain_min = 32;
  // CHM_FCHAIN_MAXD: blocks above it run as GEMM tiles. 64 by default: blocks of 65-80
  // fit the fused kernel (one workgroup per compute unit, 158 KB of LDS) but its operator
  // loads are then exposed; at L = 32, D = 2000 the tiles take 41.5 ms at the middle site
  // against the fused kernel's 46.0
  int fchain_maxd = 64;
  DBuf<DevCBlock> cblocks_f[FCHAIN_NB];
  int ncblocks_f[FCHAIN_NB] = {};
  DBuf<DevCBlock> cblocks_l;  // blocks of the large variant (same groups and terms)
  int ncblocks_l = 0;
  int chain_P_l = 1;
  bool chain_large_ok = false;  // the device grants the large variant its LDS
  DBuf<DevBRed> creds;          // output blocks split over several work items
  int ncreds = 0;
  DBuf<double> cpart;
  DBuf<DevPair> d_pairs;
  DBuf<DevKTerm> d_kterms;
  HostStage stage;
  // Heff preparation: pair groups and term buffers kept across sites
  struct PGrp { int out, in; int64_t q0, q1, elems; };
  std::vector<std::vector<PGrp>> p_grp;
  std::vector<std::vector<Term>> term_pool;
  // update executor: staging and scratch kept across calls (no page faults per call)
  HostStage ustage;
  HostStage cstage;  // chained-kernel records of the prepared site
  HostStage tstage;  // GEMM-tile records of the prepared site
  DBuf<DevCombo> cmb_dc;  // combination records (site preparation and operator updates)
  DBuf<DevSrc> cmb_ds;
  DBuf<int64_t> cmb_cs;
  DBuf<double> u_arena;   // combinations of the operator update
  std::vector<int> u_qof, u_seg;
  std::vector<int64_t> u_woff, u_order, u_hist;
  DBuf<DevGemmOp> u_d1, u_d2;
  DBuf<DevTile> u_dt1, u_dt2;
  DBuf<double> u_W;
  DBuf<DevOItem> u_it1, u_it2;
  DBuf<DevPRed> u_rd;
  DBuf<double> u_part;
  bool use_ops_kernel = false;  // CHM_OPS_KERNEL=1: small update products through k_ops, others on tiles
  int ops_maxd = OPS_MAXD;     // CHM_OPS_MAXD: largest m, n routed to k_ops (at most OPS_MAXD)
  double prep_s[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // seconds of the last prepare, by phase
  bool verbose = false;                         // CHM_VERBOSE: phase times on stderr
  int64_t item_elems = 32768;        // CHM_ITEM_ELEMS: matrix elements per work item
  DBuf<DevBItem> bitems;
  DBuf<DevBRed> breds;
  DBuf<double> bpart;
  int nbitems = 0, nbreds = 0;

  // optional timing of the three kernels of one application (events on the library stream)
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_ms[3] = {0.0, 0.0, 0.0};
  int64_t t_count = 0;

  // MPS site tensors on the device (layouts from the current sectors)
  std::vector<TLay> mps;
  std::vector<double*> mps_buf;
  double last_update_flops = 0.0;
  int64_t last_update_terms = 0;

  // two-site object on the device (program convention) and the sweep driver
  DBuf<double> Sdev;
  int32_t s_buf = -1;
  int s_site = -1;
  void* rocblas = nullptr;  // rocblas_handle for the SVDs of Split
  static constexpr int SVD_STREAMS = 4;  // side streams of the large Split SVDs
  int64_t jacobi_retries = 0;            // Jacobi sectors decomposed again by rocSOLVER (not converged)
  int svd_max_sweeps = 0;                // CHM_SVD_MAX_SWEEPS: sweep limit of the Jacobi kernel (0: its own)
  std::vector<hipStream_t> svd_streams;
  std::vector<void*> svd_handles;
  std::vector<hipEvent_t> svd_ev_out;
  hipEvent_t svd_ev_in = nullptr;
  void side_streams(int n);
  int svd_jacobi_maxn = SVD_JACOBI_MAXN;  // CHM_SVD_JACOBI_MAXN: larger centre sectors use rocSOLVER
  double total_min = 1e8, max_disc_last = 0.0, last_min = 1e8;
  double tm[6] = {0, 0, 0, 0, 0, 0};  // seconds: join, prepare, davidson, split, update, total
  // Heff applications inside sweeps: F_H summed over applications and their device time
  // (events around each application, read back after every Davidson)
  double mv_flops = 0.0, mv_ms = 0.0;
  bool mv_timing = false;
  std::vector<hipEvent_t> mv_ev;
  int mv_used = 0;
  void mv_flush();
  int64_t n_matvec_total = 0;
  // one record per two-site step of the sweeps since chm_dmrg_init: site, energy, Heff
  // applications, Davidson seconds, F_H, V and the device ms of its Heff applications
  struct SiteRec {
    int site, products;
    double energy, seconds, F_H, V, heff_ms;
  };
  std::vector<SiteRec> site_log;
  uint64_t noise_seed = 1;
  void join(int site);
  // Sobject::Join of MPS tensors Tl, Tr laid out on sectors s into buffer id out_buf
  // (layout Sj); the terms run through run_terms.
  void join_terms(const Sectors& s, const TLay& Tl, const TLay& Tr, int st, const SLayout& Sj, int32_t out_buf);
  // sites first..last in the sweep direction (-1: the end of the chain); a range that
  // does not start at the chain end continues the sweep begun by the previous range
  double sweep(bool moving_right, int D, double rtol, double noise_prefactor, bool change, int first = -1,
               int last = -1);
  double sweep_noise = 0.0;
  double split(int site, int D, bool moving_right, bool change);
  void mps_random(uint64_t seed);
  void presolve();
  double solve_site(int index, double rtol, double noise, int D, bool moving_right, bool change, int plan_next = -1);
  double dmrg_solve(int n, const int* D, const double* econv, const int* max_sweeps, const double* noise, const double* rtol);

  ~Ctx();
  void mps_alloc(int site);
  OpT* find_tensor(int cnt, int kind, int k, int m);
  void update_operators(int index, bool mr);
  // Executes operator-update terms: every block in outs is overwritten with the sum of
  // its terms (zero if none).
  void run_terms(const std::vector<USeg>& segs, const std::vector<std::pair<Ref, std::pair<int, int>>>& outs,
                 const double* combo_base);
  void run_terms(const std::vector<UTerm>& terms, const std::vector<std::pair<Ref, std::pair<int, int>>>& outs,
                 const double* combo_base);
  void set_problem(int L, int group, const int* irr, int N, int TwoS, int Irrep, const double* mx, double econst);
  const double* resolve_with(const Ref& r, const double* combo_base) const;
  void set_sectors(int b, int n, const int* Ns, const int* Ss, const int* Is, const int* ds);
  OpT* add_tensor(int cnt, bool mr, int kind, int k, int m, int two_j, int n_elec, int n_irrep, int64_t n);
  void clear_set(int cnt);
  void prepare(int site);
  const double* resolve(const Ref& r) const;
  void apply(const double* dxp, double* dyp);
  void apply_with(const double* xp, double* yp, double* bp, double* cp, double* lp, double* ld);
  int64_t bpart_need = 0, cpart_need = 0;  // partial-sum doubles the prepared site's product uses
  // frees the prepared site's explicit matrix and tile scratch (and the update arena,
  // which the next update allocates again); the site must be prepared again before the
  // next product
  void release_site_scratch() {
    Mvals.release();
    Mcol.release();
    bpart.release();
    W.release();
    u_arena.release();
    M_nnz = 0;
    site = -1;
  }
  void diag(double* d);
  // Davidson with its scalars on the device and one iteration captured as a HIP graph
  // (davidson); the host-driven form (davidson_host) serves the learned preconditioner,
  // sample harvesting and the per-kernel timing mode. CHM_DAVIDSON_DEVICE=0 selects the
  // host-driven form everywhere, CHM_DAVIDSON_GRAPH=0 runs the device form without graphs.
  double davidson(double* x, double rtol, int* n_matvec);
  double davidson_host(double* x, double rtol, int* n_matvec);
  bool dav_device = true, dav_graph = true;
  DBuf<double*> dav_vec, dav_hv;
  DBuf<const double*> dav_ptrs;
  DBuf<double> dav_coef, dav_part;
  DBuf<DevDav> dav_state;
  DevDav* dav_host = nullptr;  // page-locked copy of the scalars after each iteration
  double wall_khz = 1e5;       // device wall clock (wall_clock64) ticks per ms
  int64_t dav_deflations = 0, dav_graphs = 0, dav_graph_launches = 0;
  void symm(double* x, bool to_symm);
};

void auto_sectors(HostState& c, int D);
std::vector<OpShape> opset_shapes(const HostState& c, int cnt, bool mr);

}  // namespace chm

// The opaque handle of the C ABI.
struct chm_ctx {
  chm::Ctx c;
};
