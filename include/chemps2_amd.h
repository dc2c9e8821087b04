/*
 * chemps2_amd: MI355X (gfx950) implementation of CheMPS2's two-site DMRG
 * effective-Hamiltonian path, behind a C ABI.
 *
 * The library replaces, for one two-site step at sites (i, i+1), the work the
 * reference does in Heff::SolveDAVIDSON (CheMPS2/Heff.cpp:317-386): the matrix-vector
 * product Heff::makeHeff (Heff.cpp:43-248 with HeffDiagrams1-5.cpp), the diagonal
 * Heff::fillHeffDiag (Heff.cpp:250-315, HeffDiagonal.cpp) and the Davidson solver
 * (Davidson.cpp:31-536). The caller keeps the reference's bookkeeping and operator
 * construction and hands over, per step, the data that Heff reads: the L^4 integral
 * table, the sector dimensions and the renormalised operator tensors.
 *
 * Conventions
 *  - All entry points return 0 on success and a negative CHM_ERR_* code otherwise;
 *    chm_last_error() describes the last failure. Nothing throws across the ABI.
 *  - Dense blocks are column-major FP64, in exactly the storage order of the
 *    reference objects they replace (TensorOperator.cpp:29-102, Sobject.cpp:36-147).
 *  - Pointers are host pointers unless an on_device flag says otherwise; device
 *    pointers must be valid on the context's device and hold FP64 data of the stated
 *    length. The library reads and writes them on its own stream (chm_stream), which
 *    is not ordered against the caller's streams: the caller finishes the work that
 *    writes an input before the call (or makes chm_stream wait on it), and reads an
 *    output only after the call returned (every on_device call synchronises its stream
 *    before returning).
 *  - A context is used from one host thread at a time.
 */
#ifndef CHEMPS2_AMD_H
#define CHEMPS2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chm_ctx chm_ctx;

#define CHM_OK 0
#define CHM_ERR_ARG (-1)
#define CHM_ERR_STATE (-2)
#define CHM_ERR_MISSING (-3)
#define CHM_ERR_SIZE (-4)
#define CHM_ERR_HIP (-5)
#define CHM_ERR_NUMERIC (-6)

/* Operator kinds, the arrays DMRG keeps per boundary (CheMPS2/include/chemps2/DMRG.h:
 * Ltensors, F0tensors, F1tensors, S0tensors, S1tensors, Atensors, Btensors, Ctensors,
 * Dtensors, Qtensors, Xtensors; allocation in DMRGoperators.cpp:909-1145). */
enum chm_kind {
  CHM_L = 0, CHM_F0 = 1, CHM_F1 = 2, CHM_S0 = 3, CHM_S1 = 4,
  CHM_A = 5, CHM_B = 6, CHM_C = 7, CHM_D = 8, CHM_Q = 9, CHM_X = 10
};

/* Create / destroy a context bound to HIP device `device` with its own stream. */
int chm_ctx_create(int device, chm_ctx** out);
int chm_ctx_destroy(chm_ctx* ctx);
const char* chm_last_error(const chm_ctx* ctx);

/* The problem in DMRG orbital order: replaces Problem (Problem.h:44: L, N, 2S, target
 * irrep, orbital irreps via gIrrep) and Problem::gMxElement (Problem.cpp:351-384).
 * mx: the L^4 table h_{ab;cd} at mx[a + L(b + L(c + L d))]. group: 0..7 as Irreps.h. */
int chm_set_problem(chm_ctx* ctx, int L, int group, const int* orb_irreps, int N, int two_s, int irrep,
                    const double* mx, double econst);

/* Current virtual dimensions of boundary b (0..L): replaces SyBookkeeper::gCurrentDim
 * (SyBookkeeper.cpp:165). Sectors not listed have dimension 0. Replaces any previous
 * list for b. */
int chm_set_sectors(chm_ctx* ctx, int boundary, int n, const int* N, const int* two_s, const int* irrep,
                    const int* dim);

/* Dimensions of every boundary derived as SyBookkeeper(Prob, D) does (FCI dimensions,
 * capped and scaled to D; SyBookkeeper.cpp:29-269). */
int chm_set_sectors_auto(chm_ctx* ctx, int D);
int chm_get_sectors(const chm_ctx* ctx, int boundary, int* n, int* N, int* two_s, int* irrep, int* dim);

/* The problem set by chm_set_problem: Problem::gL, gN, gTwoS, gIrrep and the number of
 * irreps of the group (Problem.h:44-80, Irreps::getNumberOfIrreps). */
int chm_get_problem(const chm_ctx* ctx, int* L, int* N, int* two_s, int* irrep, int* nirrep);

/* One operator tensor of slot cnt (boundary cnt+1), built moving right (1) or left
 * (0): the storage of TensorOperator::gStorage() (TensorOperator.cpp:29-102) for
 * DMRG::<kind>tensors[cnt][k][m] (k, m ignored where the array has fewer indices).
 * two_j, n_elec, n_irrep as TensorOperator::get_2j/get_nelec/get_irrep. n must equal
 * the tensor's size for the current sectors. data is a host pointer. */
int chm_tensor_upload(chm_ctx* ctx, int cnt, int moving_right, int kind, int k, int m, int two_j, int n_elec,
                      int n_irrep, const double* data, int64_t n);
/* Allocates every operator of slot cnt with the reference's shapes
 * (DMRGoperators.cpp:909-1145) and fills them with deterministic pseudo-random values
 * (benchmark inputs). */
int chm_tensor_fill_synthetic(chm_ctx* ctx, int cnt, int moving_right, uint64_t seed);
/* Drops the operators of slot cnt (DMRG::deleteTensors). */
int chm_tensor_clear(chm_ctx* ctx, int cnt);

/* MPS site tensor `site` in the storage of TensorT (TensorT.cpp:38-104; blocks
 * (NL, 2SL, IL) x (NR, 2SR, IR) for the current sectors of boundaries site and site+1). */
int chm_mps_upload(chm_ctx* ctx, int site, const double* data, int64_t n);
int chm_mps_download(chm_ctx* ctx, int site, double* data, int64_t n);
/* Number of doubles of site's TensorT storage for the current sectors
 * (TensorT::gKappa2index(gNKappa()), TensorT.cpp:38-104). */
int chm_mps_layout_size(const chm_ctx* ctx, int site, int64_t* n);
/* Copy of one operator tensor of slot cnt to the host (the TensorOperator storage). */
int chm_tensor_download(chm_ctx* ctx, int cnt, int kind, int k, int m, double* data, int64_t n);
/* Renormalised operators of slot `index` built on the device from the neighbouring slot
 * and one MPS tensor: DMRG::updateMovingRight(index) (moving_right = 1, uses site
 * `index`, slot index-1) or updateMovingLeft(index) (site index+1, slot index+1)
 * (DMRGoperators.cpp:243-907 with TensorL/F0/F1/S0/S1/Q/X and
 * TensorOperator::update). */
int chm_update_operators(chm_ctx* ctx, int index, int moving_right);

/* Sobject::Join (Sobject.cpp:212-258) of the device MPS tensors of site and site+1
 * into the context's two-site object (program convention); V receives its length. */
int chm_join(chm_ctx* ctx, int site, int64_t* V);
/* Read / overwrite the context's two-site object (program convention). */
int chm_s_get(chm_ctx* ctx, double* x, int64_t n, int on_device);
int chm_s_set(chm_ctx* ctx, const double* x, int64_t n, int on_device);
/* Sobject::Split (Sobject.cpp:260-622) of the two-site object into the MPS tensors of
 * site and site+1 with at most D states (change = 0 keeps the dimensions); needs the
 * FCI dimensions of chm_set_sectors_auto. The discarded weight goes to *disc_weight. */
int chm_split(chm_ctx* ctx, int site, int D, int moving_right, int change, double* disc_weight);

/* The whole DMRG on the device, mirroring CheMPS2::DMRG (DMRG.h:93-103): init =
 * SyBookkeeper(Prob, D) dimensions, a random left-normalised MPS from `seed`
 * (TensorT::random + left_normalize) and PreSolve (DMRG.cpp:255-265); solve = the
 * ConvergenceScheme instructions (D, energy tolerance, maximum sweeps, noise
 * prefactor, Davidson tolerance) swept as DMRG::Solve (DMRG.cpp:268-417); energy =
 * the lowest energy met (TotalMinEnergy, with the constant of the Hamiltonian). */
int chm_dmrg_init(chm_ctx* ctx, int D, uint64_t seed);
/* DMRG::PreSolve (DMRG.cpp:255-265) on the current MPS (for instance one loaded from a
 * checkpoint with chm_mps_upload): operators of every boundary rebuilt moving right,
 * TotalMinEnergy and the last sweep's discarded weight reset. Needs the FCI dimensions
 * of chm_set_sectors_auto for the later Splits. */
int chm_dmrg_presolve(chm_ctx* ctx);
int chm_dmrg_solve(chm_ctx* ctx, int n, const int* D, const double* econv, const int* max_sweeps, const double* noise,
                   const double* rtol, double* energy);
/* Excited states (DMRG::activateExcitations / newExcitation, DMRG.cpp:464-505): the
 * current MPS and its virtual dimensions become the next lower state, with level shift
 * eshift (> 0); a new state starts from SyBookkeeper(Prob, D) dimensions and a random
 * MPS from `seed`, and PreSolve builds its operators together with the overlap operators
 * TensorO to every lower state (TensorO.cpp:39-193, DMRGoperators.cpp:555-567, 888-900).
 * Every later chm_dmrg_solve / sweep then adds, per two-site step, the projector
 * sum_q |VeffTilde_q><VeffTilde_q| of the lower states (DMRG::calcVeffTilde,
 * DMRGtechnics.cpp:540-620) to Heff, so that it converges to the next state
 * (H + sum_q eshift_q |q><q|). The lower states stay in HBM until chm_dmrg_init.
 * chm_dmrg_nstates: the number of states (lower states + 1). */
int chm_dmrg_new_excitation(chm_ctx* ctx, double eshift, int D, uint64_t seed);
int chm_dmrg_nstates(const chm_ctx* ctx, int* n);
/* The 2-RDM of the current state, DMRG::calc_rdms_and_correlations without the 3-RDM
 * and correlations (DMRGtechnics.cpp:40-110; TwoDM, TwoDM.cpp:40-640): after a solve
 * (the state a right sweep leaves), the MPS is brought into left-canonical form and one
 * left sweep fills the spin-summed 2-RDMs site by site (TwoDM::FillSite) with
 * right-normalisation and left-moving operator updates. A and B (host arrays of L^4
 * doubles, or null) receive TwoDM's two_rdm_A / two_rdm_B in DMRG orbital order, element
 * (i, j, k, l) at i + L (j + L (k + L l)), divided by 2S + 1 as
 * TwoDM::correct_higher_multiplicities does; *energy (or null) = Econst + 0.5 sum A h
 * (TwoDM::energy). Afterwards the MPS is right-canonical and the operators are those of
 * a finished left sweep (the next sweep moves right, or call chm_dmrg_presolve after a
 * left-normalising sweep). */
int chm_dmrg_2rdm(chm_ctx* ctx, double* A, double* B, double* energy);
/* One half sweep, DMRG::sweepleft (moving_right = 0) or DMRG::sweepright
 * (moving_right = 1) of DMRG.cpp:354-417 with one instruction's D, Davidson tolerance
 * and noise prefactor; change = 0 keeps the virtual dimensions (the reference's first
 * sweep). *energy: the energy at the last site. */
int chm_dmrg_sweep(chm_ctx* ctx, int moving_right, int D, double rtol, double noise, int change, double* energy);
/* Sites first..last of one sweep in its direction (moving left: L-2 >= first >= last >= 1;
 * moving right: 0 <= first <= last <= L-3); a range that does not start at the chain end
 * continues the sweep of the previous range (noise level and minimum energy). Splits
 * DMRG::sweepleft / sweepright (DMRG.cpp:357-417) at a site, for callers that act between
 * the steps (checkpoints, checks). */
int chm_dmrg_sweep_range(chm_ctx* ctx, int moving_right, int first, int last, int D, double rtol, double noise,
                         int change, double* energy);
/* Heff applications inside the sweeps since chm_dmrg_init: the sum of F_H over them
 * and their device time in ms (HIP events around each application). No counterpart
 * in the reference: it is the sweep-averaged Heff.x rate of the benchmark. */
int chm_dmrg_heff_stats(const chm_ctx* ctx, double* flops, double* kernel_ms);
/* Per two-site step of the sweeps since chm_dmrg_init (or the last read): with null
 * arrays *n receives the number of records; otherwise sites[n], products[n] (Heff
 * applications of its Davidson) and values[5 n] (energy, Davidson seconds, F_H, V,
 * device ms of its Heff applications) are written and the records cleared. The
 * reference prints the same per site (Heff.cpp:376, DMRG.cpp:441-445). */
int chm_dmrg_site_log(chm_ctx* ctx, int* n, int* sites, int* products, double* values);
/* seconds[6]: join, Heff preparation, Davidson, split, operator updates, all sweeps;
 * n_matvec: Heff applications. */
int chm_dmrg_stats(const chm_ctx* ctx, double* seconds, int64_t* n_matvec);

/* Two-site step at (site, site+1): builds the work lists of Heff (the reference's
 * Heff(denBK, Prob, rtol) + Sobject(site, denBK), Heff.h:50, Sobject.cpp:36-147).
 * Needs slot site-1 built moving right (site > 0) and slot site+1 moving left
 * (site < L-2). */
int chm_heff_prepare(chm_ctx* ctx, int site);
/* Davidson vector length V = Sobject::gKappa2index(gNKappa()). */
int chm_heff_veclen(const chm_ctx* ctx, int64_t* V);
/* y = Heff x in the symmetric convention (Heff::makeHeff, Heff.cpp:43-248). */
int chm_heff_apply(chm_ctx* ctx, const double* x, double* y, int on_device);
/* chm_heff_apply (on_device = 1) for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. x and y are device arrays of V doubles; the partial sums
 * of the product live in a static device array of the library (256 MiB per GPU, held
 * until the process ends) instead of the context's buffers, and the call makes no
 * hipMalloc / hipFree. A prepared site whose partial sums exceed the array is refused
 * with CHM_ERR_SIZE. The array is shared by every call on a GPU: calls that overlap on
 * one GPU race. Returns after the product is complete on the context's stream. */
int chm_heff_apply_noalloc(chm_ctx* ctx, const double* x, double* y);
/* Lower states for excited-state DMRG at the prepared site: n_lower vectors of length V
 * (row-major, n_lower x V, in the block order and symmetric convention of the x of
 * chm_heff_apply), each the reference's VeffTilde
 * of one lower state (DMRG::calcVeffTilde, DMRGtechnics.cpp:540-620: the lower state's
 * two-site tensor contracted with the overlap operators, already scaled by
 * sqrt(E_shift)). chm_heff_apply then adds sum_q (v_q . x) v_q
 * (Heff::addDiagramExcitations, HeffDiagrams1.cpp:65-86) and chm_heff_diag adds
 * sum_q v_q^2 (Heff::addDiagonalExcitations, HeffDiagonal.cpp:621-642); chm_davidson
 * uses both. The vectors are copied; n_lower = 0 removes them; chm_heff_prepare clears
 * them. */
int chm_heff_set_lower(chm_ctx* ctx, int n_lower, const double* veff, int on_device);
/* Heff forms (no reference counterpart; the reference computes in FP64 only).
 * CHM_FORM_FP64: every product in FP64. CHM_FORM_MIXED: the explicit sector matrix is
 * stored as float pairs (hi = fl32(v), lo = fl32(v - hi)); chm_davidson iterates with
 * the hi floats until the residual norm falls below max(rtol, 1e-5), then restarts
 * from the Ritz vector with hi + lo products (v to 2^-48) until it falls below rtol.
 * Products and sums are FP64 in both forms. Takes effect at the next chm_heff_prepare. */
#define CHM_FORM_FP64 0
#define CHM_FORM_MIXED 1
int chm_set_heff_form(chm_ctx* ctx, int form);
/* y = Heff x with the hi floats only (a site prepared in CHM_FORM_MIXED); chm_heff_apply
 * is the refined product in that form. */
int chm_heff_apply_f32(chm_ctx* ctx, const double* x, double* y, int on_device);
/* Products chm_davidson made with the hi floats and with full precision since the
 * context was created. */
int chm_mixed_stats(const chm_ctx* ctx, int64_t* low_products, int64_t* full_products);
/* Learned preconditioner (no reference counterpart; the reference uses the diagonal of
 * Heff, Davidson.cpp:322-350). A per-element MLP in FP32 (CHM_MLP_IN = 8 inputs: asinh(d
 * - theta), r / rms(r), u sqrt(V), log(dL) / 4, log(dR) / 4, three zeros; three hidden
 * layers of 128 with GELU; one output m) scales the Davidson correction:
 * t = -(1 + m) (r + alpha u) / (d - theta), 1 + m clamped to [0.1, 10], alpha keeping t
 * orthogonal to u. CHM_PRECOND_DIAGONAL (default) is the reference's preconditioner. */
#define CHM_PRECOND_DIAGONAL 0
#define CHM_PRECOND_LEARNED 1
#define CHM_MLP_IN 8
int chm_precond_mode(chm_ctx* ctx, int mode);
/* The learned correction is kept only while each step it makes lowers the Davidson
 * residual norm; the first step that does not switches that solve back to the plain
 * diagonal correction. fallbacks: the number of solves switched back since the context
 * was created. */
int chm_precond_stats(const chm_ctx* ctx, int64_t* fallbacks);
/* Parameters: W1 (8 x 128), b1, W2 (128 x 128), b2, W3 (128 x 128), b3, W4 (128 x 1), b4,
 * each W row-major (inputs x outputs); FP32. */
int chm_mlp_nparams(int64_t* n);
int chm_mlp_set(chm_ctx* ctx, const float* params, int64_t n);
int chm_mlp_get(chm_ctx* ctx, float* params, int64_t n);
/* out[i] = m of sample i (X: n x CHM_MLP_IN, host arrays). */
int chm_mlp_forward(chm_ctx* ctx, const float* X, int64_t n, float* out);
/* loss = mean_i w_i (p_i (1 + m_i) - e_i)^2 and its gradient (nparams floats). */
int chm_mlp_loss_grad(chm_ctx* ctx, const float* X, const float* p, const float* e, const float* w, int64_t n,
                      float* grad, double* loss);
/* Adam (lr, 0.9, 0.999, 1e-8) over shuffled batches; mean batch loss of the first and last
 * epochs. */
int chm_mlp_train(chm_ctx* ctx, const float* X, const float* p, const float* e, const float* w, int64_t n, int batch,
                  int epochs, double lr, uint64_t seed, double* loss_first, double* loss_last);
/* Training samples from chm_davidson (diagonal preconditioner only): per_iteration
 * elements of every preconditioner application, with p the plain preconditioned
 * residual, e the corresponding element of (x* - u) for the converged x* divided by the
 * application's best scale (p.e)/(p.p) (applications where it is not positive are
 * dropped), and w 1 / sum e^2 of that application. chm_harvest_take with null arrays
 * returns the count; with arrays it copies and clears. */
int chm_harvest(chm_ctx* ctx, int per_iteration);
int chm_harvest_take(chm_ctx* ctx, int64_t* n, float* X, float* p, float* e, float* w);
/* d = diag(Heff) (Heff::fillHeffDiag, Heff.cpp:250-315; nLower = 0). */
int chm_heff_diag(chm_ctx* ctx, double* d, int on_device);
/* Lowest eigenpair of Heff from the guess x (program convention, as Sobject storage
 * before Heff::SolveDAVIDSON; Heff.cpp:330-386 with Davidson(V, 32, 3, rtol, 1e-12)).
 * On return x holds the eigenvector in the program convention. */
int chm_davidson(chm_ctx* ctx, double* x, int on_device, double rtol, double* eigenvalue, int* n_matvec);
/* Counters of the Davidson solves of this context: out[0] deflations (restarts at 32
 * vectors, Davidson.cpp:352-412), out[1] iteration graphs captured, out[2] graph launches
 * (one per iteration after the first of a solve). Measurement; no reference counterpart. */
int chm_davidson_stats(const chm_ctx* ctx, int64_t* out);

/* Routing and work-item sizes of the Heff kernels (no reference counterpart): the
 * largest |kappa||kappa'| of the explicit sector matrix, terms per chained work item and
 * per input group, the MFMA chained kernel for blocks up to 48 (1) or the LDS one (0),
 * and fchain_min: chained blocks whose largest dimension is above it (and at most 64, or
 * CHM_FCHAIN_MAXD up to 80) run in the fused MFMA kernel k_fchain, the rest of those up
 * to 64 in the LDS kernel; a negative value (or 0 for the two term counts) keeps the
 * current setting.
 * chm_set_heff_knobs sets the defaults. chm_set_tuning installs the autotuner's table:
 * n rows of 9 values (largest block dimension lo, hi; V lo, hi; then the five knobs as
 * above); chm_heff_prepare uses the first row whose ranges hold the site's largest
 * block dimension and V, or the defaults; a knob given in the environment when the
 * context was created (CHM_EXPLICIT_MAX, CHM_CHAIN_ITEM_TERMS, CHM_CHAIN_GROUP_TERMS,
 * CHM_CHAIN_MFMA, CHM_FCHAIN_MIN) keeps that value. chm_tuning_used: out[7] = row used (-1: none),
 * largest block dimension, then the five knobs of the last chm_heff_prepare. */
int chm_set_heff_knobs(chm_ctx* ctx, int64_t explicit_max, int item_terms, int group_terms, int chain_mfma,
                       int fchain_min);
int chm_set_tuning(chm_ctx* ctx, int n, const int64_t* rows);
int chm_tuning_used(const chm_ctx* ctx, int64_t* out);
/* F_H: 2mnk summed over every dgemm_ the reference's makeHeff issues at this site;
 * B_H: 8 * (2V + sum of the sizes of the operator tensors it reads), in bytes. */
int chm_flops(const chm_ctx* ctx, double* F_H, double* B_H);
/* stats[16]: terms, combinations, combination doubles, W doubles, chain phase-1 tiles,
 * chain phase-2 tiles, explicit block pairs, chain terms, blocks, V, diagonal terms,
 * explicit matrix doubles, explicit column indices, explicit size threshold,
 * explicit work items, 0. */
int chm_plan_stats(const chm_ctx* ctx, int64_t* stats);
/* The context's HIP stream (hipStream_t) for callers that time or order work on it. */
int chm_stream(const chm_ctx* ctx, void** stream);
/* Timing of chm_heff_apply with HIP events on the context stream, accumulated per
 * kernel group: ms[0] GEMM tiles of large chained blocks (both phases), ms[1] explicit
 * block product, ms[2] output-stationary chained kernels; count = number of
 * applications timed. */
int chm_timing_enable(chm_ctx* ctx, int enable);
int chm_timing_read(const chm_ctx* ctx, double* ms, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* CHEMPS2_AMD_H */
