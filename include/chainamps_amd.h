/*
 * chainamps_amd.h -- C ABI of the MI355X-native ChainAMPS hot path.
 *
 * The library keeps QN-blocked tensors resident on one MI355X (gfx950) and
 * runs the two-vertex effective-Hamiltonian matvec, its Lanczos eigensolve,
 * the blockwise SVD truncation and the iTEBD gate application with
 * hand-written HIP kernels.  Every entry point returns 0 on success and a
 * non-zero status otherwise; cam_last_error() gives the message (per
 * thread).  Plain pointers and sizes only; complex numbers are interleaved
 * (re, im) doubles exactly as std::complex<double>.
 *
 * Tensor conventions follow the reference MPX_matrix: an ordered list of
 * directed indices, the first `num_row_indices` of them forming the rows
 * of a CSC matrix in mixed radix (first index fastest), the rest the
 * columns (MPX_matrix.hpp:61-66, sparse_interface.cpp:85-91).
 *
 * Threading: one context per host thread; a context is not re-entrant.
 * Calls taking host buffers synchronise before returning.  A blockwise SVD
 * with two or more large blocks runs part of them on a helper thread with a
 * second stream and rocBLAS handle of the same context (default priority),
 * joined before the call returns; CAM_SVD_SERIAL=1 keeps one stream.
 */
#ifndef CHAINAMPS_AMD_H
#define CHAINAMPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cam_ctx_s cam_ctx;
typedef struct cam_tensor_s cam_tensor;
typedef struct cam_heff_s cam_heff;

/* One tensor index: MPXIndex (MPXIndex.hpp:17-65).  `charges` holds
 * size*nrules shorts, entry-major. */
typedef struct {
  int32_t inward;          /* 1 = Ingoing(), 0 = Outgoing() */
  int32_t physical;        /* 1 = physical (model basis) index */
  int64_t size;
  const int16_t* charges;
} cam_index_desc;

enum { CAM_MEM_HOST = 0, CAM_MEM_DEVICE = 1 };
enum { CAM_FP64 = 0, CAM_FP32 = 1 };

/* ---- context --------------------------------------------------------
 * Replaces nothing in the reference (which is single-process CPU code); a
 * context owns one HIP stream, a rocBLAS/rocSOLVER handle and scratch. */
int cam_ctx_create(int device, cam_ctx** out);
int cam_ctx_destroy(cam_ctx* ctx);
int cam_ctx_sync(cam_ctx* ctx);
/* Frees the context's grow-only workspaces (the Lanczos basis and the two
 * H_eff intermediate buffers; the next call that needs them allocates them
 * again), e.g. before a call that needs the memory for something else. */
int cam_ctx_release_workspaces(cam_ctx* ctx);
/* The hipStream_t the context launches on (for events/profiling). */
int cam_ctx_stream(cam_ctx* ctx, void** stream);
const char* cam_last_error(void);
/* Library build/version string. */
const char* cam_version(void);

/* ---- tensors (MPX_matrix storage, MPX_matrix.cpp:29-35) ------------------
 * Upload a CSC tensor; the zero-flux rule IngoingRowState(r) ==
 * OutgoingColState(c) (MPX_matrix.cpp:275-302) is enforced: a violating
 * entry is an error. `rules` are the Z_N charge rules (states.hpp:32-34). */
int cam_tensor_from_csc(cam_ctx* ctx, int nrules, const int16_t* rules,
                        int nidx, const cam_index_desc* idx, int num_row_indices,
                        int64_t rows, int64_t cols, const int64_t* p,
                        const int64_t* i, const double* x, cam_tensor** out);
/* Number of indices / row indices, and one index descriptor (charges
 * copied to `charges` when non-NULL). */
int cam_tensor_shape(const cam_tensor* t, int* nidx, int* num_row_indices, int* nrules);
int cam_tensor_index(const cam_tensor* t, int k, int32_t* inward, int32_t* physical,
                     int64_t* size, int16_t* charges);
/* Stored block-sparse element count (upper bound of the CSC nnz). */
int cam_tensor_nnz(const cam_tensor* t, int64_t* nnz);
/* Download as CSC (exact zeros dropped); p has cols+1 entries, i and x
 * capacity `cap` entries; *nnz receives the count. */
int cam_tensor_to_csc(const cam_tensor* t, int64_t cap, int64_t* p, int64_t* i,
                      double* x, int64_t* nnz);
int cam_tensor_destroy(cam_tensor* t);
/* MPX_matrix::ShiftNumRowIndices (MPX_matrix.cpp:72-85) -- metadata only
 * in block storage. */
int cam_tensor_shift_num_row_indices(cam_tensor* t, int num_row_indices);
/* MPX_matrix::RemoveDummyIndices (MPX_matrix.cpp:415-457). */
int cam_tensor_remove_dummy_indices(cam_tensor* t, int n, const int64_t* which);
/* MPX_matrix::MoveDummyIndices (MPX_matrix.cpp:87-105). */
int cam_tensor_move_dummy_indices(cam_tensor* t, int n, const int64_t* order);
/* Scale all values (MPX_matrix::Rescale, MPX_matrix.cpp:459-462). */
int cam_tensor_rescale(cam_tensor* t, double re, double im);

/* ---- contraction ------------------------------------------------------
 * ajaj::contract (MPX_matrix.hpp:25, MPX_matrix.cpp:744-791): output
 * indices are A's kept indices (flipped if conjA) then B's kept indices
 * (flipped if conjB); num_row_indices = #kept A.  `pairs` holds npairs
 * (a,b) index pairs.  One grouped MFMA launch over all QN-compatible block
 * pairs. */
int cam_contract(cam_ctx* ctx, const cam_tensor* A, int conjA, const cam_tensor* B,
                 int conjB, int npairs, const int64_t* pairs, cam_tensor** out);
/* ajaj::reorder (MPX_matrix.cpp:630-638). */
int cam_reorder(cam_ctx* ctx, const cam_tensor* A, int conjA, int n,
                const int64_t* order, int num_row_indices, cam_tensor** out);

/* ---- two-vertex effective Hamiltonian --------------------------------
 * TwoVertexComponents (DMRG_routines.hpp:189-214, DMRG_routines.cpp:779-833)
 * from the left block L (6 indices), the MPO W (4) and the right block R (6).
 * *n receives the length of the QN-allowed compressed vector. */
int cam_heff_create(cam_ctx* ctx, const cam_tensor* L, const cam_tensor* W,
                    const cam_tensor* R, cam_heff** out, int64_t* n);
int cam_heff_destroy(cam_heff* h);
/* allowed_indices of DMRG_routines.cpp:808-824 (loop order a2>s2>a0>s1):
 * n flat positions into the d*chiL*d*chiR vector. */
int cam_heff_allowed_map(const cam_heff* h, int64_t* idx);
/* Excited-state projector terms of TwoVertexMPOMPSMultiply
 * (DMRG_routines.cpp:947-973; TwoVertexComponents' ProjectorTensors,
 * DMRG_routines.cpp:781-806): every later matvec adds, for each of the
 * nproj tensors P[k] (host, n complex in the reference compressed order,
 * entries of the projector MPX_matrix at allowed_indices) with weight
 * weights[k], weights[k] * (sum_i P[k][i] in[i]) * conj(P[k]) -- the
 * contraction with the vector without conjugation, as contract00112233 --
 * in FP64 and FP32 matvecs alike.  With projectors the eigensolver never
 * takes the dense route (as DMRG_routines.cpp:852).  nproj = 0 removes
 * them; at most 64. */
int cam_heff_set_projectors(cam_heff* h, int nproj, const double* const* P, const double* weights);
/* Algorithmic flop count of one matvec (sum of 8*m*n*k over the
 * QN-allowed block triples of the four-stage order L, W, W, R). */
int cam_heff_flops(const cam_heff* h, double* flops);
/* TwoVertexMPOMPSMultiply (DMRG_routines.cpp:898-974): out = H_eff . in,
 * in/out of length n in the reference compressed order.  mem_kind selects
 * host or device buffers; precision CAM_FP64 or CAM_FP32 (split-complex
 * MFMA, FP32 accumulation). */
int cam_heff_apply(cam_heff* h, const double* in, double* out, int mem_kind, int precision);
/* Same matvec on vectors already in the device block layout (the Lanczos
 * layout), no gather/scatter; `in`/`out` device pointers of n complex. */
int cam_heff_apply_blocked(cam_heff* h, const double* in, double* out, int precision);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 cam_heff_apply_blocked with the matvec's two
 * intermediate buffers in a per-device scratch array of
 * CAM_NOALLOC_SCRATCH_BYTES (256 MiB, held by every GPU the library runs on
 * until the process ends) instead of the context workspaces.  It calls no
 * hipMalloc, hipFree or allocating library routine; it fails with an error
 * when the intermediates exceed the array (slice the matvec with
 * cam_heff_set_slices first) or when the stage plans are not yet on the
 * device at their launch configuration (one cam_heff_apply_blocked call, or
 * cam_heff_set_launch, puts them there).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race. */
#define CAM_NOALLOC_SCRATCH_BYTES (256LL << 20)
int cam_heff_apply_blocked_noalloc(cam_heff* h, const double* in, double* out);
/* The FP32 matvec on FP32 split-complex vectors in the device block layout
 * (n complex = 2n floats each), as the mixed-precision eigensolver runs it:
 * no FP64 conversion of in or out. */
int cam_heff_apply_blocked32(cam_heff* h, const float* in, float* out);
/* Benchmark helper (no reference counterpart): run the matvec `reps` times
 * on internal random block-layout vectors and return, per stage
 * (L, W, W, R), the average launch duration in ms measured with HIP events
 * on the context stream and the algorithmic flops of that stage. */
int cam_heff_profile(cam_heff* h, int precision, int reps, double* stage_ms, double* stage_flops);
/* The same on the caller's device block-layout vectors (FP64, or FP32
 * split complex for CAM_FP32), for sizes where internal vectors do not fit
 * next to the caller's. */
int cam_heff_profile_device(cam_heff* h, int precision, int reps, const void* in, void* out, double* stage_ms,
                            double* stage_flops);
/* Sliced matvec (no reference counterpart: TwoVertexMPOMPSMultiply,
 * DMRG_routines.cpp:898-974, forms LeftPart.V whole).  The a0' rows of
 * every sector are cut into row_slices ranges and the a2 columns into
 * col_slices; each slice runs the four stages on intermediates of
 * 1/(row_slices col_slices) the size, and the column slices accumulate
 * into the output.  Results are the same up to the order of the a2 sum.
 * By default cam_heff_create takes the fewest slices whose two
 * intermediate buffers fit the free device memory (plus the context's
 * existing workspaces, less 4 GB): 1 x 1 wherever the whole intermediates
 * fit.  set_slices fixes the counts; set_workspace_limit takes the fewest
 * (powers of two) whose buffers fit `bytes` at `precision` (FP32
 * intermediates take half the FP64 bytes); get_slices reports the counts
 * and the FP64 workspace bytes.  The workspaces are sized at the first
 * matvec of each precision.  The adjoint needs 1 x 1. */
int cam_heff_set_slices(cam_heff* h, int row_slices, int col_slices);
int cam_heff_set_workspace_limit(cam_heff* h, double bytes, int precision);
int cam_heff_get_slices(const cam_heff* h, int* row_slices, int* col_slices, double* workspace_bytes);
/* Algorithmic HBM bytes per stage (L, W, W, R) of the FP64 matvec: every
 * operand read once and the output written once (psi -> T1 -> T2 -> T3 ->
 * out, plus the environment or MPO blocks), 16 bytes per element. */
int cam_heff_stage_bytes(const cam_heff* h, double* stage_bytes);
/* Launch configuration of the four stage GEMMs (no reference counterpart:
 * the reference's SpGEMM has no tunables).  By default the first matvec at
 * each precision autotunes them (timed candidates, cached per block-size
 * histogram; CAM_AUTOTUNE=0 disables).  set fixes it for every stage:
 * algo 0 = four real MFMAs per complex product, 1 = three (Gauss);
 * tile_mode 0 = 64x64 tiles, 1 = per-block least padding, 2 = 32x64/64x32
 * only, 3 = least padding including 128x64 / 64x128 / 128x32 tiles; wg_per_cu =
 * workgroups per CU in the persistent grid; kchunk = k per LDS stage (16,
 * or 8 for half the LDS and more resident workgroups).
 * get returns {algo, tile_mode, wg_per_cu, kchunk} per stage (16 ints). */
int cam_heff_set_launch(cam_heff* h, int precision, int algo, int tile_mode, int wg_per_cu, int kchunk);
int cam_heff_get_launch(const cam_heff* h, int precision, int* cfg);
/* Reference order <-> device block layout (K3 gather/scatter). */
int cam_heff_to_blocked(cam_heff* h, const double* in_dev, double* out_dev);
int cam_heff_from_blocked(cam_heff* h, const double* in_dev, double* out_dev);

/* Lowest eigenpair ("SR"), replacing arpack_eigs<TwoVertexComponents,...>
 * (arpack_interface.hpp:85-111, 174-237) at TwoVertexComponents::HED
 * (DMRG_routines.cpp:835-896).  guess: host vector (reference order) or
 * NULL.  converge=0 gives the reference's growth-step setting (tol 0.1);
 * converge=1 tolerance `tol` (<=0: machine precision).  ncv: Krylov size
 * (reference 20).  mixed=1 runs the bulk in FP32 with FP64 residual
 * correction and final FP64 steps.  evec: host, reference order, phase
 * gauge of DMRG_routines.cpp:870-890 applied.  As in the reference, an
 * H_eff of dimension n <= SPARSE_THRESHOLD (1000) is instead formed densely
 * and diagonalised (DMRG_routines.cpp:850-858): guess, tol, ncv and mixed
 * are unused, the eigenvector has entries below machine epsilon cut and no
 * phase gauge.  info (may be NULL) receives {matvecs_fp64, matvecs_fp32,
 * restarts, residual_norm}. */
int cam_heff_lowest(cam_heff* h, const double* guess, int converge, double tol,
                    int ncv, int mixed, double* eval, double* evec, double* info);
/* The same eigensolve on a device vector in the block layout (the form the
 * iDMRG driver keeps between steps; no host round trip of the n-vector):
 * x holds the start vector when have_start (a zero vector means a random
 * start) and receives the eigenvector. */
int cam_heff_lowest_blocked(cam_heff* h, double* x, int have_start, int converge, double tol,
                            int ncv, int mixed, double* eval, double* info);

/* ---- blockwise SVD truncation -------------------------------------------
 * MPX_matrix::SVD -> SparseMatrix::SVD (MPX_matrix.cpp:157-192,
 * sparse_interface.cpp:1248-1406): per QN column block SVD in FP64 on the
 * GPU, global descending sort (ties by block then index), keep
 * min(D, #values) (D=0: all), tail truncation below max_trunc, degeneracy
 * extension (relative gap < 1e-3), U/V^dagger entries below eps/rows
 * dropped, new bond charges from U's rows.  values has capacity cap;
 * *nvals receives the kept count; *discarded the discarded weight. */
int cam_svd_truncate(cam_ctx* ctx, const cam_tensor* theta, int64_t D, double max_trunc,
                     cam_tensor** U, cam_tensor** Vd, double* values, int64_t cap,
                     int64_t* nvals, double* discarded);

/* ---- iTEBD gate ------------------------------------------------------------
 * iTEBD::apply_and_decompose (TEBD_routines.cpp:233-266): A0, A1 left-shape
 * MPS tensors, G the 4-index bond operator, lambda0 (n0 values) the bond
 * weights on A0's left bond.  Produces the new A0, A1 (swapped as in the
 * reference), lambda1 (capacity cap) and the truncation. */
int cam_tebd_gate(cam_ctx* ctx, const cam_tensor* G, const cam_tensor* A0,
                  const cam_tensor* A1, const double* lambda0, int64_t n0, int64_t D,
                  double minS, cam_tensor** A0n, cam_tensor** A1n, double* lambda1,
                  int64_t cap, int64_t* n1, double* truncation);

/* ---- gate construction ---------------------------------------------------
 * MPO_matrix::ExtractMPOBlock (MPO_matrix.cpp:69-85): rows [r0, r1] of the
 * left MPO bond, columns [c0, c1] of the right one. */
int cam_extract_mpo_block(cam_ctx* ctx, const cam_tensor* W, int64_t r0, int64_t r1, int64_t c0,
                          int64_t c1, cam_tensor** out);
/* MakeOddBondHamiltonian (odd=1) / MakeEvenBondHamiltonian (odd=0)
 * (TEBD_routines.cpp:826-867): [s1', s2', s1, s2, dummy, dummy], two row
 * indices. */
int cam_bond_hamiltonian(cam_ctx* ctx, const cam_tensor* W, int odd, cam_tensor** out);
/* MakeBondEvolutionOperator (TEBD_routines.cpp:869-880): exp(-i timestep H)
 * from the blockwise Hermitian decomposition (Exponentiate,
 * sparse_interface.cpp:1707-1714; SparseMatrix::HED 1414-1441); the result
 * keeps the first four indices of H. */
int cam_bond_evolution_operator(cam_ctx* ctx, const cam_tensor* H, double timestep, cam_tensor** out);

/* ---- iDMRG growth --------------------------------------------------------
 * iDMRG::run / SuperBlock::grow_two_vertex (iDMRG_routines.cpp:19-33,
 * DMRG_routines.cpp:309-374) with every tensor on the device: W the model
 * MPO, target the target charges (nrules shorts).  converge=0 is the
 * reference growth-step eigensolver setting (tol 0.1); mixed=1 runs the
 * bulk of each eigensolve in FP32.  Each cam_idmrg_grow call adds two
 * vertices; record[CAM_IDMRG_RECORD] receives {N, energy, energy/N,
 * entropy, truncation, chi kept, n, matvecs fp64, matvecs fp32, residual,
 * seconds, heff flops, eigensolve seconds, SVD seconds, block update +
 * H_eff build + prediction seconds, 1 - fidelity of the prediction
 * (CheckConvergence, DMRG_routines.cpp:769-771, the Energies.dat column),
 * 1 when no prediction was made for the next step (the new bond's two
 * sides carry different charges, i.e. a non-zero target; that step then
 * starts at random), vertices in the left block, vertices in the right
 * block}. */
#define CAM_IDMRG_RECORD 19
typedef struct cam_idmrg_s cam_idmrg;
int cam_idmrg_create(cam_ctx* ctx, const cam_tensor* W, const int16_t* target, int converge,
                     int mixed, int ncv, cam_idmrg** out);
int cam_idmrg_grow(cam_idmrg* d, int64_t chi, double truncation, double* record);
/* ---- finite DMRG sweeps ---------------------------------------------------
 * One finite step on the grown chain: SuperBlock::move_right_two_vertex
 * (to_right=1) or move_left_two_vertex (DMRG_routines.cpp:376-444).  The
 * vertex leaving the centre joins the block on that side (MakeLeftBlock /
 * MakeRightBlock, or the initial block at an end), the prediction is
 * MakeRFinitePrediction / MakeLFinitePrediction (709-716) from the stored
 * MPS matrix of the vertex that enters, the eigensolve uses the object's
 * converge / mixed / ncv setting, the SVD keeps chi states with the
 * truncation switched off next to an end (-0.0), and the record's fidelity
 * field is 1 - |<prediction|U s Vd>| (CheckConvergence(Prediction,
 * MPSDecomposition), 773-776).  Blocks and MPS matrices of every position
 * stay on the device (the reference's .BLOCK / .MPS_matrix files).  Growth
 * after a finite step is an error. */
int cam_idmrg_move(cam_idmrg* d, int to_right, int64_t chi, double truncation, double* record);
/* FiniteDMRG::run (fDMRG_routines.cpp:19-66): num_sweeps sweeps, each
 * moving right to the end, left to the other end and right back to the
 * middle; records receives CAM_IDMRG_RECORD doubles per step and midpoint
 * (optional) 1 for the steps the reference writes to Energies.dat (centre
 * in the middle, on the way left and on the way back).  With records NULL
 * only *nrec (the number of steps) is set. */
int cam_fdmrg_sweeps(cam_idmrg* d, int64_t num_sweeps, int64_t chi, double truncation, double* records,
                     int32_t* midpoint, int64_t cap, int64_t* nrec);
/* the left and right blocks of the current centre (new handles sharing the
 * device tensors; destroy them with cam_tensor_destroy) */
int cam_idmrg_blocks(const cam_idmrg* d, cam_tensor** LB, cam_tensor** RB);
/* singular values of the current central decomposition (capacity cap) */
int cam_idmrg_values(const cam_idmrg* d, double* values, int64_t cap, int64_t* n);
int cam_idmrg_destroy(cam_idmrg* d);

/* ---- iTEBD time evolution ------------------------------------------------
 * iTEBD (TEBD_routines.cpp:14-85 TrotterDecomposition, 120-185 evolve,
 * 233-266 apply_and_decompose) with the unit cell on the device: W the
 * model MPO, A0/A1 the left-shape cell tensors [s, a_l, a_r], lambda0/1
 * the bond weights, dt the time step, order 1, 2 or 4.  evolve runs
 * nsteps full time steps with no measurement in between; the second-order
 * step is odd(dt/2) even(dt) odd(dt) ... even(dt) odd(dt/2) with the odd
 * full-step operator between steps; the fourth-order run applies piece 0
 * once, then per step pieces 1..9 of Suzuki's product with
 * OrthogonaliseInversionSymmetric (measurement.cpp:131-178) after each
 * even-bond piece, then the odd full step joining two steps (as the
 * reference, the last step's closing odd half step stays applied).  record[7] receives {steps so far,
 * time, truncation of the last gate, entropy of lambda0, chi, seconds,
 * gate-contraction flops of the call}. */
typedef struct cam_itebd_s cam_itebd;
int cam_itebd_create(cam_ctx* ctx, const cam_tensor* W, const cam_tensor* A0, const cam_tensor* A1,
                     const double* lambda0, int64_t n0, const double* lambda1, int64_t n1, double dt,
                     int order, cam_itebd** out);
int cam_itebd_evolve(cam_itebd* it, int64_t nsteps, int64_t D, double minS, double* record);
/* bond weights k (0 or 1), capacity cap; *n receives the count */
/* iTEBD::evolve with measurement steps (TEBD_routines.cpp:120-231): every
 * `interval` steps of the object's step count the reference's measuring
 * branch runs (order 2: odd(dt/2), cell swap, canonisation, odd(dt/2);
 * order 4: piece 10, swap, canonisation, piece 0), the cell is canonised
 * with OrthogonaliseInversionSymmetric and do_measurements (99-118) records
 * 7 + 2 nops doubles: {step, time, truncation of the last gate, entropy of
 * lambda0, |overlap|, overlap (re, im)} with the overlap Overlap(initial
 * cell, cell) (measurement.cpp:282-288, a bra != ket transfer-matrix
 * eigenvalue), then OneVertexMeasurement (317-324) of each of the nops MPO
 * operators ops[k] as (re, im).  *nrec receives the number of records; up
 * to cap are written. */
int cam_itebd_evolve_measured(cam_itebd* it, int64_t nsteps, int64_t D, double minS, int64_t interval, int nops,
                              const cam_tensor* const* ops, double* records, int64_t cap, int64_t* nrec);
int cam_itebd_lambda(const cam_itebd* it, int k, double* values, int64_t cap, int64_t* n);
/* a copy of cell tensor k (0 or 1) as a new tensor handle */
int cam_itebd_tensor(const cam_itebd* it, int k, cam_tensor** out);
int cam_itebd_destroy(cam_itebd* it);

/* ---- finite-chain TEBD ----------------------------------------------------
 * TEBD (TEBD_routines.cpp:273-791) with the MPS in device memory instead of
 * the "Evolving_<name>_Left_v" / "_Right_k" files: sites[0..nsites-1] is a
 * left-canonical MPS (FiniteMPS::makeLC) of an even number of vertices,
 * each [s (physical), a_l (in), a_r (out)]; W the lower-triangular model
 * MPO; order 1, 2 or 4 (TrotterDecomposition, 14-85).  evolve runs nsteps
 * time steps of apply_to_odd_bonds / apply_to_even_bonds / left_canonise
 * with bond dimension D and truncation minS, measuring every `interval`
 * steps (left_canonise_measure without MultiVertexMeasurements).  records
 * returns 7 doubles per measurement: {step, time, largest truncation of the
 * step, entropy of bond N/2, overlap with the initial state (re, im),
 * |overlap|}, for up to cap records.  site returns a copy of left-canonical
 * site v (1..N). */
typedef struct cam_ftebd_s cam_ftebd;
int cam_ftebd_create(cam_ctx* ctx, const cam_tensor* W, const cam_tensor* const* sites, int64_t nsites,
                     double dt, int order, cam_ftebd** out);
int cam_ftebd_evolve(cam_ftebd* t, int64_t nsteps, int64_t D, double minS, int64_t interval);
int cam_ftebd_records(const cam_ftebd* t, double* records, int64_t cap, int64_t* n);
int cam_ftebd_site(const cam_ftebd* t, int64_t v, cam_tensor** out);
int cam_ftebd_destroy(cam_ftebd* t);

/* ---- on-disk formats (checkpoint / resume interop) --------------------------
 * The reference's binary files.  cam_tensor_store / cam_tensor_load:
 * MPX_matrix::store / load_MPX_matrix (MPX_matrix.cpp:219-247, 594-612) for
 * .MPX_matrix and .MPS_matrix files; a physical index carries no charges on
 * disk, so the loader takes the basis charges (nrules rules, d states,
 * charges d x nrules).  The unit cell: UnitCell::store / load_UnitCell_binary
 * (UnitCell.cpp:70-187), i.e. the basis (Basis::fprint_binary, states.cpp:
 * 119-189: rules, charges, energies), the MPS matrices and the lambdas;
 * loaded cells are read back through the accessors. */
int cam_tensor_store(const cam_tensor* t, const char* path);
int cam_tensor_load(cam_ctx* ctx, const char* path, int nrules, const int16_t* rules, int64_t d,
                    const int16_t* phys_charges, cam_tensor** out);
typedef struct cam_unitcell_s cam_unitcell;
int cam_unitcell_store(const char* path, int nrules, const int16_t* rules, int64_t d, const int16_t* phys_charges,
                       const double* energies, int n, const cam_tensor* const* matrices, const double* const* lambdas,
                       const int64_t* lambda_sizes);
int cam_unitcell_load(cam_ctx* ctx, const char* path, cam_unitcell** out);
/* rules (nrules entries), charges (d x nrules) and energies (d) with capacity cap states */
int cam_unitcell_basis(const cam_unitcell* u, int* nrules, int16_t* rules, int64_t* d, int16_t* charges,
                       double* energies, int64_t cap);
int cam_unitcell_size(const cam_unitcell* u, int* n);
int cam_unitcell_matrix(const cam_unitcell* u, int k, cam_tensor** out);
int cam_unitcell_lambda(const cam_unitcell* u, int k, double* values, int64_t cap, int64_t* n);
int cam_unitcell_destroy(cam_unitcell* u);

/* ---- unit cells, orthogonalisation and the energy per vertex -------------
 * cam_unitcell_create: UnitCell(spectrum, matrices, lambdas) (UnitCell.hpp:41)
 * over device tensors [s (physical), a_l (in), a_r (out)] (references, not
 * copies); the basis energies of a cell built on the device are zero.
 * cam_unitcell_from_decomposition: UnitCell(MPSDecomposition, PreviousLambda)
 * (UnitCell.hpp:42-48): [U, diag(values) . Vd . diag(previous)^-1 reordered
 * (phys, in, out)], lambdas [previous, values].
 * cam_idmrg_unitcell: the same from an iDMRG object's central decomposition
 * and previous lambda (iDMRG_DRV.cpp:70, getCentralDecomposition /
 * getPreviousLambda).
 * cam_transfer_apply: LeftComponentsMultiply (side 0) / RightComponentsMultiply
 * (side 1) (measurement.cpp:582-629) of the cell's transfer map on host
 * vectors in and out of the allowed entries in the reference's order
 * (allowed_indices, measurement.cpp:431-442: pairs (r, c) of equal charges,
 * c outer, r inner), complex as (re, im) pairs.
 * cam_transfer_leading: TransferMatrixComponents(cell, 1, Identity).LeftED
 * (side 0) or .RightED (side 1) with NUMEVALS = 1 and LARGESTMAGNITUDE
 * (measurement.cpp:405-629): the leading eigenvalue value[2] (re, im) of the
 * cell's transfer map, by a device Krylov-Schur iteration on the Hermitian
 * matrices of the boundary space (Krylov dimension ncv, <= 0 for ARPACK's
 * 20; at most 50 restarts; Ritz residual <= tol |value|, tol <= 0 for
 * 1e-14); info[4] = {matvecs, restarts, relative residual, converged}.  The
 * two boundary indices of the cell must carry the same charges (cells of a
 * zero-target iDMRG and of iTEBD do).
 * cam_orthogonalise: Orthogonalise(const UnitCell&) (measurement.cpp:
 * 180-246) of a two-vertex cell: a new cell [U^H X A0 / sqrt(l), A1 X^-1 U]
 * with lambdas [new boundary values (unit square sum), the input's second
 * lambda]; an error (the reference's empty cell) when a leading eigenvalue
 * has |Im| >= 100 eps |Re| or SqrtDR meets eigenvalues of both signs.
 * info[12] = {left value re, im, right value re, im, entropy, new bond
 * dimension, left matvecs, right matvecs, left residual, right residual,
 * kept SqrtDR vectors (left + right), seconds}.
 * cam_orthogonalise_inversion_symmetric: OrthogonaliseInversionSymmetric
 * (measurement.cpp:131-178): as cam_orthogonalise with only the left fixed
 * point, reused on the right boundary, and the SVD of X . diag(lambda0) . Y
 * (lambda0 = the input cell's first lambda); info as cam_orthogonalise (the
 * right value and matvecs repeat the left ones).
 * cam_itwovertex_energy: iTwoVertexEnergy(H_MPO, Ortho) (measurement.cpp:
 * 371-397) of an orthogonalised two-vertex cell: energy[2] (re, im) per
 * vertex; terms[6] = (on-vertex, bond inside the cell, bond to the next
 * cell) as (re, im) pairs, before the factor 1/2. */
int cam_unitcell_create(cam_ctx* ctx, int n, const cam_tensor* const* matrices, const double* const* lambdas,
                        const int64_t* lambda_sizes, cam_unitcell** out);
int cam_unitcell_from_decomposition(cam_ctx* ctx, const cam_tensor* U, const double* values, int64_t nvalues,
                                    const cam_tensor* Vd, const double* previous, int64_t nprevious,
                                    cam_unitcell** out);
int cam_idmrg_unitcell(cam_idmrg* g, cam_unitcell** out);
int cam_transfer_apply(const cam_unitcell* u, int side, const double* in, double* out);
int cam_transfer_leading(const cam_unitcell* u, int side, double tol, int ncv, double* value, double* info);
int cam_orthogonalise(const cam_unitcell* u, cam_unitcell** out, double* info);
int cam_orthogonalise_inversion_symmetric(const cam_unitcell* u, cam_unitcell** out, double* info);
int cam_itwovertex_energy(const cam_tensor* W, const cam_unitcell* u, double* energy, double* terms);

/* ---- adjoint (no counterpart in the reference) --------------------------
 * Vector-Jacobian product of E = <psi|H_eff|psi>/<psi|psi>:
 * dpsi (host, n complex, reference order) = dE/dconj(psi) and
 * dW (host, complex, same CSC positions as W's stored entries in
 * cam_tensor_to_csc order, count nnzW) = dE/dW.  Environments are
 * recomputed, never stored per Lanczos step. */
int cam_heff_vjp(cam_heff* h, const double* psi, double* dpsi, double* dW, int64_t nnzW);

/* dE/d(MPO couplings) (no counterpart in the reference): the MPO is taken
 * linear in couplings theta_c, W = sum_c theta_c G_c (+ constant entries),
 * and each generator G_c = dW/dtheta_c is given as a sparse list over W's
 * stored entries (cam_tensor_to_csc order): entries
 * gen_entry[gen_ptr[c] .. gen_ptr[c+1]) with complex coefficients gen_coef
 * (interleaved re, im).  dE[2c], dE[2c+1] = sum_k dE/dW[entry k] * coef k
 * (re, im; the imaginary part vanishes for a Hermitian MPO and is returned
 * as a check).  dpsi (host, n complex, reference order) receives dE/dconj(psi)
 * when not NULL.  Same cost as cam_heff_vjp. */
int cam_heff_coupling_grad(cam_heff* h, const double* psi, int64_t nnzW, int ncoup, const int64_t* gen_ptr,
                           const int64_t* gen_entry, const double* gen_coef, double* dE, double* dpsi);

#ifdef __cplusplus
}
#endif
#endif
