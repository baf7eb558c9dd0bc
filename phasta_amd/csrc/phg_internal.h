// Internal declarations of libphgpu: the device-resident context and the
// host-side launchers of the HIP kernels.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "phgpu.h"

void phg_set_error(const char* fmt, ...);
// PHG_EINVAL with a message when the context is matrix-free (no lhsK)
int require_matrix(phg_ctx* c, const char* who);

#define PHG_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      phg_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                    hipGetErrorString(e_));                               \
      return PHG_EHIP;                                                    \
    }                                                                     \
  } while (0)

#define PHG_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_) return rc_;       \
  } while (0)

// Rows of the device block-CSR matrix are grouped by 64 (one wavefront);
// a group holds S slots, each slot one 5x5 block per row of the group.
constexpr int kGroup = 64;

// Position of entry e (0..24, column-major 5x5) of row r's block inside one
// slot of 25 * kGroup values of type T: entries 0..23 form 16-byte vectors
// of V = 16 / sizeof(T) entries, lane-interleaved (vector k of row r at
// (k * kGroup + r) * V), so a wavefront reads each vector with one
// 16-byte load per lane over a contiguous 1 KiB; entry 24 follows as a
// plane of kGroup values.
template <typename T>
__host__ __device__ constexpr int slot_pos(int e, int r) {
  return e < 24 ? ((e / (16 / (int)sizeof(T))) * kGroup + r) * (16 / (int)sizeof(T)) +
                      e % (16 / (int)sizeof(T))
                : 24 * kGroup + r;
}

// Assembled blocks in CSR order (the reference's lhsK(25,nnz)), each padded
// to a multiple of 16 bytes (28 floats, 26 doubles) so that one lane moves
// its block with 16-byte vector accesses; the pad entries are zero.
template <typename T>
__host__ __device__ constexpr int aos_stride() { return sizeof(T) == 4 ? 28 : 26; }

// The 25 entries of row r's block in one slot (blk = slot base), read and
// written with 16-byte vector accesses (6 float4 or 12 double2) and one
// scalar for entry 24.  The matvec streams the matrix once per product and
// it does not fit in the caches, so the reads are non-temporal: they do not
// displace the gathered vector from L2 and the MALL.
typedef float phg_f4 __attribute__((ext_vector_type(4)));
typedef double phg_d2 __attribute__((ext_vector_type(2)));
template <typename A>
__device__ inline void load_slot(const float* blk, int r, A a[25]) {
  const phg_f4* v = reinterpret_cast<const phg_f4*>(blk);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const phg_f4 t = __builtin_nontemporal_load(v + k * kGroup + r);
    a[4 * k] = (A)t.x;
    a[4 * k + 1] = (A)t.y;
    a[4 * k + 2] = (A)t.z;
    a[4 * k + 3] = (A)t.w;
  }
  a[24] = (A)__builtin_nontemporal_load(blk + 24 * kGroup + r);
}
template <typename A>
__device__ inline void load_slot(const double* blk, int r, A a[25]) {
  const phg_d2* v = reinterpret_cast<const phg_d2*>(blk);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const phg_d2 t = __builtin_nontemporal_load(v + k * kGroup + r);
    a[2 * k] = (A)t.x;
    a[2 * k + 1] = (A)t.y;
  }
  a[24] = (A)__builtin_nontemporal_load(blk + 24 * kGroup + r);
}
template <typename A>
__device__ inline void store_slot(float* blk, int r, const A a[25]) {
  float4* v = reinterpret_cast<float4*>(blk);
#pragma unroll
  for (int k = 0; k < 6; ++k)
    v[k * kGroup + r] = make_float4((float)a[4 * k], (float)a[4 * k + 1], (float)a[4 * k + 2],
                                    (float)a[4 * k + 3]);
  blk[24 * kGroup + r] = (float)a[24];
}
template <typename A>
__device__ inline void store_slot(double* blk, int r, const A a[25]) {
  double2* v = reinterpret_cast<double2*>(blk);
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k * kGroup + r] = make_double2((double)a[2 * k], (double)a[2 * k + 1]);
  blk[24 * kGroup + r] = (double)a[24];
}

// Largest number of element blocks one assembly lane sums (the blocks of
// one CSR entry are split into chunks of at most this size).  Measured at
// 128^3 (float / FP64 block pass): 2: 7.90 / 12.74 ms, 3: 7.81 / 12.41,
// 4: 8.07 / 12.77, 6: 8.28 / 12.92, 9: 8.51 / 13.23; shorter chunks give
// more lanes and even loop trips for the 24 blocks of a diagonal entry and
// the 4 to 6 of an off-diagonal one.
constexpr int kChunk = 3;
// Physical constants and time-integration factors used in the kernels.
struct PhysParams {
  double Rgas, gamma, gamma1, pr, mu;
  double dtsfct, taucfct, temper;
  double almi, alfi, gami, Dtgl;
  int itau;  // 0 Shakib's diagonal tau, 1 Franca's; 10 and 11 the matrix tau
};

// Per-node Dirichlet mask: bit f set = equation/dof f (local order
// p,u,v,w,T) is constrained (bc3lhs.f, bc3res.f, bc3bdg.f).
__host__ __device__ inline uint8_t constraint_mask(int ibc) {
  uint8_t m = 0;
  if (ibc & 4) m |= 1;                         // pressure
  if (((ibc >> 3) & 7) == 7) m |= 2 | 4 | 8;   // no-slip velocity
  if (ibc & 2) m |= 16;                        // temperature
  return m;
}

// Learned preconditioner state (k_mpnn.hip); allocated by phg_mpnn_init.
struct MpNet {
  bool ready = false;
  int use_pc = 0;            // FGMRES: directions per cycle preconditioned
                             // by the network (0: block-diagonal GMRES)
  bool keep_pair = false;    // keep (raw BDiag, b, z) of the last solve
  // resident time stepping: between phg_predict and phg_update a rejected
  // network stays set aside for the remaining Newton iterations of the step
  bool in_step = false;
  bool step_off = false;
  long long n_act = 0, nedge = 0;
  int* ptr = nullptr;        // periodic node graph, CSR over all nodes
  int* col = nullptr;
  int* rev = nullptr;        // index of the reverse edge
  float4* ev = nullptr;      // e_ij (x, y, z) per edge
  float4* esum = nullptr;    // sum_j e_ij per node
  float* inv = nullptr;      // 1/deg (0 for slaves)
  float* dscale = nullptr;   // 25 input scales of BDiag
  double* bd_raw = nullptr;  // BDiag before factorisation (25 planes)
  bool bd_raw_ok = false;    // bd_raw holds the BDiag of a solve (iprec != 0)
  const double* bd_src = nullptr;
  double* zhat = nullptr;    // scaled-system solution of the last solve
  uint16_t* x = nullptr;     // n x 32 inputs (bf16)
  uint16_t* h[4] = {};       // n x 64 hidden states (bf16)
  uint16_t* A[3] = {};       // n x 320 aggregated inputs (bf16)
  float* U[4] = {};          // n x 64 pre-activations (training)
  float* Y = nullptr;        // n x 5 outputs (training)
  float* Ystar = nullptr;    // n x 5 targets (training)
  float* dh = nullptr;       // n x 64
  float* dU = nullptr;       // n x 64
  float* dA = nullptr;       // n x 320
  float* dY = nullptr;       // n x 5
  float* sq = nullptr;       // n squared errors
  float* part = nullptr;     // weight-gradient partials
  int nparts = 0;
  float* w = nullptr;        // master weights (float)
  float* g = nullptr;        // gradient
  float* m1 = nullptr;       // Adam moments
  float* m2 = nullptr;
  uint16_t* wt[4] = {};      // bf16 weights, Wt[c][k]
  uint16_t* wr[4] = {};      // bf16 weights, W[k][c]
  double* sbuf = nullptr;    // [0] ||v||^2, [1] s, [2] loss, [3] ||r||^2, [4] loss sum,
                             // [8..11] acceptance-test dot products
  double* vin = nullptr;     // 5 planes, host-array entry points
  double* zout = nullptr;    // 5 planes
  double* rbuf = nullptr;    // 5 planes (residual training)
  double* gbuf = nullptr;    // 5 planes
  double* zbasis = nullptr;  // FGMRES preconditioned directions (K vectors)
  double* wplain = nullptr;  // K v_1 of the block-diagonal first direction (acceptance test)
  int adam_t = 0;
  bool train_ready = false;
  // parameters shared with another context (training over several meshes):
  // w, g, m1, m2, wt, wr then belong to param_owner
  phg_ctx* param_owner = nullptr;
  // device-resident training pairs: per slot 35 floats per node
  // (BDiag before factorisation 25, b 5, z 5)
  float* ds = nullptr;
  int ds_cap = 0;
};

struct phg_ctx {
  long long nshg = 0, nel = 0, nelb = 0, nnz = 0;
  int ndofBC = 0;
  int precision = PHG_FP64;
  bool matrix_free = false;     // PHG_FP64_MATRIX_FREE: no lhsK, SolMFG only
  phg_params prm{};
  hipStream_t stream = nullptr;
  long long dev_bytes = 0;

  // mesh (device)
  double* x = nullptr;          // 3 planes of nshg
  int* ienc = nullptr;          // 4 planes of nel, 0-based
  // assembly: point states of the elements and the block-owner lanes
  long long nlanes = 0;
  int4* lanes = nullptr;        // block-list offset, CSR entry, count, flags
  long long nlanes_diag = 0;
  int4* lanes_diag = nullptr;   // the lanes of the diagonal entries alone
  long long nlanes_pair = 0;
  int4* lanes_pair = nullptr;   // k_asm_pairs: pair wavefronts, then diagonal ones
  int* lanes_k2 = nullptr;      // per lane of lanes_pair: transposed entry or -1
  int* blocks = nullptr;        // e*16 + a*4 + b, grouped per owner lane
  void* scratch = nullptr;      // this device's address of the library's
                                // static staging array (phg_spsi3pre_noalloc)
  double* qps = nullptr;        // (nel*4) x 12 point records (FP64 path)
  float* qps32 = nullptr;       // (nel*4) x 12, mixed path
  double* geo = nullptr;        // nel x 14: shape gradients, WdetJ (+ pad)
  float* geo32 = nullptr;       // nel x 16
  double* resid = nullptr;      // (nel*4) x 5 element residual per node (element
                                // order), then the element products of k_jv and
                                // k_mres in the incidence slots
  double* dcw = nullptr;        // DC per point (iDC != 0): nel*4 factors, or nel*4 x 8
                                // DC giju with the matrix tau
  size_t dcw_len = 0;           // doubles allocated in dcw
  void* mtau = nullptr;         // nel*4 W PTau per point (itau = 10), the tangent's type
  int* ienbc = nullptr;         // 4 planes of nelb, colour order, 0-based
  uint8_t* bcm = nullptr;       // constraint mask per node
  int* iper = nullptr;          // 0-based master
  double* BCv = nullptr;        // (nshg, ndofBC)
  std::vector<int> colb_off;    // wall-element colour offsets (host)

  // periodic fold: masters that own slaves, slaves ascending per master
  int nfold = 0;
  int* fold_master = nullptr;
  int* fold_ptr = nullptr;
  int* fold_slave = nullptr;

  // block-CSR in groups of 64 rows ("slot" layout)
  long long ngroups = 0, nslots = 0;
  std::vector<long long> h_grp_slot;
  long long* grp_slot = nullptr;   // ngroups+1
  int* scol = nullptr;             // nslots*64 raw column (0-based)
  int* scolp = nullptr;            // nslots*64 periodic-master column
  uint32_t* tslot = nullptr;       // nslots*64 slot of the transposed block
                                   // (built on demand for training)
  uint32_t* csr2slot = nullptr;    // nnz -> slot-row index
  int* slot2csr = nullptr;         // slot-row -> CSR block index, -1 = padding
  double* lhs64 = nullptr;         // nslots*64*25, scaled matrix for the matvec
  float* lhs32 = nullptr;          // mixed precision copy

  // assembly targets, block/node contiguous (reference lhsK(25,nnz) order)
  double* acc_res = nullptr;       // 5 per node
  double* acc_bd = nullptr;        // 25 per node
  double* lhs_aos = nullptr;       // aos_stride<double>()*nnz, reference lhsK order
  float* lhs_aos32 = nullptr;      // aos_stride<float>()*nnz (mixed)
  double* nodepack = nullptr;      // 16 per node: y (local order), ac, x
  // global reconstruction of the diffusive flux (idiff 1, 3)
  long long* inc_goff = nullptr;   // per 64-node group: offset of its sliced entries
  int* inc = nullptr;              // e*4 + a per (node, incident element), sliced by
                                   // 64-node groups, element order per node, -1 padding
  int* eslot = nullptr;            // (nel*4): slot of (e, a) in the incidence layout
  double* qrec = nullptr;          // nel x 16 element records (allocated on first use)
  double* qnode = nullptr;         // 16 per node: q (12), 1/lumped mass

  // nodal vectors (SoA planes of nshg)
  double* res = nullptr;
  double* rmes = nullptr;
  double* bdiag = nullptr;         // 25 planes
  float* bdiag32 = nullptr;
  double* y = nullptr;
  double* ac = nullptr;
  double* yold = nullptr;
  double* acold = nullptr;
  double* dy = nullptr;
  double* krylov = nullptr;        // (Kspace+1) vectors of 5*nshg
  int kcap = 0;

  // reductions
  int red_blocks = 0;
  double* red_part = nullptr;      // (kcap + 3) channels of red_blocks partials
  double* scal = nullptr;          // device scalars (H column etc.)
  double* h_scal = nullptr;        // pinned mirror
  double reslast = 0.0;
  int ntotGM = 0;
  int diag_flags = 0;              // phg_diagnostics

  // scratch for host-array entry points
  double* tmp = nullptr;
  size_t tmp_bytes = 0;

  MpNet mp;

  // matrix-free GMRES (SolMFG): allocated on first use
  struct {
    bool on = false;         // gmres_cycle's product is Au1MFG
    double eg = 0.0;         // the finite-difference interval of the solve
    double* ypre = nullptr;  // U new2old(y), 5 planes
    double* t1 = nullptr;    // 5 planes
    double* t2 = nullptr;    // 5 planes
    double* pack = nullptr;  // 8 per node: the perturbed variables
  } mfg;

  // timing events on the library stream, created once
  hipEvent_t ev[3] = {};
  hipEvent_t kev[4] = {};
  hipEvent_t aev[3] = {};  // assembly: before pass 1, after pass 1, after pass 2
};

// ---- launchers (k_assembly.hip) ----
int launch_fill(phg_ctx* c, void* p, size_t bytes);
int launch_assembly(phg_ctx* c, bool lhs, bool prec);
int launch_diffusive_flux(phg_ctx* c);
int launch_bc_res_bdiag(phg_ctx* c, bool prec);
// w = A v with A the FP64 assembled tangent (after bc3LHS) of the state of
// the last assembly, v read at the periodic masters and w folded (bc3per):
// the operator of SparseAp + bc3per before the two-sided scaling.  Uses
// c->resid as scratch, so it runs after the assembly's block pass.
int launch_jv(phg_ctx* c, const double* v, double* w);

// matrix-free GMRES (SolMFG): the modified residual of the packed variables
// yp (8 per node; nullptr: y itself) at the resident state into out
// (+= with accumulate), after bc3Res; c->resid is its scratch
int launch_mres(phg_ctx* c, const double* yp, double* out, bool accumulate, bool iabs);
// ypack = U^-1 (ypre + s d) per node in the local order (d may be nullptr;
// without lu no U^-1), with itrBC's Dirichlet values and periodic copy when bc
int launch_mfg_point(phg_ctx* c, const double* ypre, const double* d, double s, double* ypack,
                     bool bc, bool lu);
// w = (L^-1 t - rmes) / e
int launch_mfg_fwd_diff(phg_ctx* c, const double* t, const double* rmes, double e, double* w);
// mode 0: u = a - (u1 - u2) / h; mode 1: u = (u1 - 2 a) / h
int launch_mfg_combine(phg_ctx* c, const double* a, const double* u1, const double* u2, double h,
                       double* u, int mode);

// ---- learned preconditioner (k_mpnn.hip) ----
int mp_param_count();
int mp_init(phg_ctx* c, const phg_mesh* m, const float* dscale);
int mp_free(phg_ctx* c);
int mp_alloc_train(phg_ctx* c);
int mp_pack(phg_ctx* c);
// z = v + y(v)/s on 5-plane vectors; train keeps the activations
int mp_forward(phg_ctx* c, const double* v, double* z, bool train);
// loss and gradient for targets ystar (n x 5, float, device) of the last
// train forward; lr > 0 applies an Adam step
int mp_backward(phg_ctx* c, float lr, double* loss_out, bool given_dy = false);
// one step on the residual loss ||K z(v) - v||^2 / ||v||^2 (v: 5 planes)
int mp_train_residual(phg_ctx* c, const double* v, float lr);
// targets (t - v) s of the last forward's v (t, v: 5 planes, device)
int mp_targets(phg_ctx* c, const double* t, const double* v);
int mp_share(phg_ctx* dst, phg_ctx* src);
int mp_dataset(phg_ctx* c, int capacity);
int mp_store_pair(phg_ctx* c, int slot);
// widen a stored pair into bd_raw, vin, zout
int mp_load_pair(phg_ctx* c, int slot);

// ---- launchers (k_solver.hip) ----
int launch_i3lu(phg_ctx* c, double* diag, double* r, int code);
int launch_spsi3pre(phg_ctx* c);
int launch_spmv(phg_ctx* c, const double* p, double* q);
int launch_sub(phg_ctx* c, double* r, const double* b, const double* q, long long n);
// q = K^T x (x overwritten by its periodic gather); needs c->tslot
int launch_spmv_t(phg_ctx* c, double* x, double* q);
// padded assembled blocks <-> lhsK(25,nnz) in double
int launch_aos_to_ref(phg_ctx* c, double* ref);
int launch_ref_to_aos(phg_ctx* c, const double* ref);
// n doubles, n even, 16-byte aligned
int launch_copy(phg_ctx* c, double* dst, const double* src, long long n);
int launch_bc3per(phg_ctx* c, double* v, int nq);
// reductions: per-block partials in channels of red_blocks doubles,
// summed in a fixed order by launch_finish (np0 partials in the first
// channel, np in the others)
int launch_dot(phg_ctx* c, const double* a, const double* b, long long n,
               double* out);
int launch_dot_part(phg_ctx* c, const double* a, const double* b, long long n,
                    int chan, int* np);
int launch_mgs_part(phg_ctx* c, double* u, const double* ujm1, const double* uj,
                    int chan_prev, int np_prev, long long n, int chan, int* np);
int launch_finish(phg_ctx* c, int chan0, int np0, int np, int nsums, double* out);
// classical Gram-Schmidt of u against nv contiguous basis vectors:
// h[0..nv-1] = (u, basis_j) of the incoming u, u -= sum h_j basis_j,
// h[nv] = (u, u) afterwards
int launch_cgs(phg_ctx* c, double* u, const double* basis, int nv, long long n, double* h);
int launch_scale(phg_ctx* c, double* u, const double* src, long long n,
                 const double* sq_ptr, double* norm_out);
int launch_combine(phg_ctx* c, double* dy, const double* basis, int k,
                   const double* coef, long long n);
int launch_stats(phg_ctx* c, double* out);
int launch_csr_to_slots(phg_ctx* c, const double* ref, double* slots);
int launch_pack_nodes(phg_ctx* c);
int launch_slots_to_csr(phg_ctx* c, const double* slots, double* ref);
int launch_slots_to_csr32(phg_ctx* c, const float* slots, double* ref);
int launch_predict(phg_ctx* c);
int launch_correct(phg_ctx* c);
int launch_update(phg_ctx* c);
int launch_itrbc(phg_ctx* c, double* y, double* ac);
// yshuffle of 5-plane vectors (src and dst may be the same)
int launch_yshuffle(phg_ctx* c, const double* src, double* dst, bool new2old);
