// Launchers of the fixed (not run-time generated) prover kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lg {

// Called by every launcher below right before its kernel launch, once the
// caller's arguments are final: the prover uses it to push the small tables it
// staged in pinned host memory to the device in one copy per launch.
using PreLaunchHook = void (*)(void* user);
void set_prelaunch_hook(PreLaunchHook fn, void* user);
// Called by every launch wrapper before it launches: runs the calling
// thread's hook (staging flush) and returns true, or returns false without
// launching anything while launches are suppressed on this thread (a prove
// that replays a recorded HIP graph re-runs its host logic only).
bool prelaunch();
void set_launch_suppression(bool on);
bool launch_suppressed();

constexpr int NTT_MAX_LOG_LDS = 14;   // one workgroup holds a size-2^14 transform in LDS
constexpr int NTT_MAX_OUT = 8;

// One transform: input column, one destination per output, optional per-job
// post-scale table applied after the inverse transform.
struct NttJob {
  const uint32_t* in;
  const uint32_t* post;
  uint32_t* out[NTT_MAX_OUT];
};

// out_o = DFT(pre_o * x) for o < n_out, where x = post * iDFT(in) when
// inverse_first is set and x = in otherwise; with n_out == 0 the inverse
// result itself is written.  Orders: inputs natural; outputs natural or
// bit-reversed per output.  Scale tables are n-entry Montgomery arrays.
enum NttKind { NTT_KIND_LDE = 0, NTT_KIND_QUOT_COSET, NTT_KIND_QUOT_BACK, NTT_KIND_CHUNKS, NTT_KIND_DEEP_LDE };

struct NttCfg {
  int kind;  // NttKind: which phase launches it (kernel symbol only)
  int logn;
  int inverse_first;
  int n_out;
  const uint32_t* pre[NTT_MAX_OUT];
  int bitrev_out[NTT_MAX_OUT];
};

// pres_dev: device array holding cfg.pre[0 .. n_out) (the host copy in cfg is
// only used for validation).
// lad_fwd / lad_inv: the radix-16 pass twiddle tables (Tables::ntt_ladders).
void launch_ntt(const NttJob* jobs_dev, int n_jobs, const NttCfg& cfg, const uint32_t* const* pres_dev,
                const uint32_t* tw_fwd, const uint32_t* tw_inv, const uint32_t* lad_fwd, const uint32_t* lad_inv,
                hipStream_t s);

// Two-pass (four-step) transform for 2^15 <= n <= 2^22, n = n1 n2 with
// n1 = 2^ceil(logn/2): out = post * DFT(pre * in) (natural input; natural or
// bit-reversed output); tmp: n-word scratch per job.  tw_a / tw_b: w^k
// (k < size/2) of the size-n2 / size-n1 sub-transforms; wlo: w_n^k (k < 2048),
// whi: w_n^{2048 k} (k < n/2048); all in the direction of the transform.
struct BigJob {
  const uint32_t* in;
  uint32_t* out;
  uint32_t* tmp;
  const uint32_t* pre;
  const uint32_t* post;
};
void launch_ntt_big(const BigJob* jobs_a, const BigJob* jobs_b, int n_jobs, int logn, const uint32_t* tw_a,
                    const uint32_t* tw_b, const uint32_t* wlo, const uint32_t* whi, int bitrev_out, hipStream_t s);
// The two passes alone (pass A: column transforms of size 2^b, twiddle,
// transpose into tmp; pass B: row transforms of size 2^a of tmp into out).
void launch_ntt_pass_a(const BigJob* jobs, int n_jobs, int a, int b, const uint32_t* tw_a, const uint32_t* wlo,
                       const uint32_t* whi, hipStream_t s);
void launch_ntt_pass_b(const BigJob* jobs, int n_jobs, int a, int b, const uint32_t* tw_b, int bitrev_out,
                       hipStream_t s);

// Fused middle pass of a large LDE (n = 2^(a+b) >= 2^15, a = ceil(log n / 2)):
// the inverse transform's row pass and the first pass of every coset's
// forward transform, with the factors swapped.  For 8 rows k2 of tmp
// (tmp[k2 2^a + j1], the inverse's pass-A output) a workgroup takes the
// inverse DFT over j1 (giving the coefficients c[k2 + 2^b k1]), and for each
// output o: scales them by pre_o (geometric: pre_o[k2] * blk_o[k1 position]),
// takes the forward DFT over k1, multiplies by w_n^(k2 i1) and stores
// u_o[i1 2^b + k2], the input of pass B with the factors swapped.
struct LdeMidJob {
  const uint32_t* tmp;               // 2^(a+b) words
  uint32_t* u[NTT_MAX_OUT];          // per output, 2^(a+b) words
};
struct LdeMidArgs {
  int a, b, n_out;
  const uint32_t* pre[NTT_MAX_OUT];  // per output: pre_o[k] for k < 2^b (its first 2^b entries)
  const uint32_t* blk[NTT_MAX_OUT];  // per output: blk_o[p] = s_o^(2^b brev_a(p)), p < 2^a
  const uint32_t* lad_inv;           // radix-16 pass ladders (Tables::ntt_ladders)
  const uint32_t* lad_fwd;
  const uint32_t* tw_inv;            // w_{2^a}^-e, e < 2^(a-1) (r-stage roots)
  const uint32_t* tw_fwd;            // w_{2^a}^e
  const uint32_t* wlo;               // w_n^e, e < 2048
  const uint32_t* whi;               // w_n^(2048 e), e < n / 2048
};
void launch_lde_mid(const LdeMidJob* jobs, int n_jobs, const LdeMidArgs& args, hipStream_t s);

// Leaf hashing: for instance b (grid y) and pair leaf k, hashes the message
// [LEAF, hdr1, hdr2, ...] whose element m >= 3 is column msgmap[m] >> 1 at row
// 2k + (msgmap[m] & 1); columns are colptrs[b * ncols + c].  Digest (8 words)
// to outs[b] + 8k.
void launch_leaf_hash(const uint32_t* const* colptrs, int ncols, const int32_t* msgmap, int msglen,
                      uint32_t hdr1, uint32_t hdr2, uint32_t* const* outs, int nleaves, int batch,
                      hipStream_t s, int threads = 256);

// Merkle level: nodes[w + i] = H(nodes[2(w+i)], nodes[2(w+i)+1]) for i < w,
// for every tree trees[b] (flat 1-indexed node array, 8 words per node).
void launch_merkle_level(uint32_t* const* trees, int width, int batch, hipStream_t s, int threads = 256);
// Levels below `width_hi` down to `width_lo` (inclusive) in one pass over
// workgroup-local subtrees (width_hi / width_lo <= 256).
void launch_merkle_levels_fused(uint32_t* const* trees, int width_hi, int width_lo, int batch, hipStream_t s);
// Same with one octet of lanes per node (cooperative Poseidon2, DPP sums):
// ~6x lower latency per level, for levels with few nodes per batch
// (width_hi / width_lo <= 32).
void launch_merkle_coop(uint32_t* const* trees, int width_hi, int width_lo, int batch, hipStream_t s);
// The top of a tree (width_hi <= 32 nodes down to the root, one workgroup per
// instance) followed by the FRI fold challenge of the root, as
// launch_fri_challenge computes it.
void launch_merkle_coop_challenge(uint32_t* const* trees, int width_hi, const uint32_t* prefix, int n_prefix,
                                  const uint32_t* prev, uint32_t* dig, uint32_t* alpha, uint32_t* roots, int batch,
                                  hipStream_t s);
// roots[8b .. 8b+8) = trees[b][8 .. 16)  (node 1 of every tree)
void launch_copy_roots(const uint32_t* const* trees, uint32_t* roots, int batch, hipStream_t s);
// Injection fold: nodes[w + j] = H(nodes[w + j], inj[b][j]).
// count injections at one level width, applied in order (inj[k * batch + b]).
void launch_inject_seq(uint32_t* const* trees, const uint32_t* const* inj, int count, int width, int batch,
                       hipStream_t s);
void launch_inject(uint32_t* const* trees, const uint32_t* const* inj, int width, int batch, hipStream_t s);

// Batched copy: for job k, dst[k][0 .. n) = src + off[k] (n words each).
void launch_scatter(const uint32_t* src, const uint64_t* off, uint32_t* const* dst, int n, int n_jobs, hipStream_t s);
// dst[k][j] = src[k][j] for j < n, k < n_jobs (device-resident inputs).
void launch_copy_cols(const uint32_t* const* src, uint32_t* const* dst, int n, int n_jobs, hipStream_t s);

// Grand product: gp[0] = 1, gp[i] = gp[i-1] * num[i-1] / den[i-1]; num, den,
// ratio (scratch) and gp are E6 columns as 6 limb arrays (limb stride n).
// Scans of n >= SCAN_MIN_TILED rows run over many workgroups and need
// scan_tmp_words(n, batch) words of scratch (scan_tmp); smaller ones need none.
constexpr int SCAN_MIN_TILED = 1 << 13;
size_t scan_tmp_words(int n, int batch);
// The allocation-free grand product (noalloc.hip): noalloc_init looks up the
// device's file-scope scratch array once (device current); the variant runs
// lg_grand_product's kernels with all scratch inside that array.
uint32_t* noalloc_init(int device);
void grand_product_noalloc(int device, hipStream_t s, const uint32_t* num, const uint32_t* den, uint32_t* gp,
                           size_t n);
// Issue-rate probe on `device`: lane operations per ns of v_add_u32,
// v_mul_lo_u32, v_mad_u64_u32 and the Montgomery product kb_mul (gops[0..3]),
// each in eight independent chains per lane over 16 workgroups per CU, and
// the device's clock (MHz).
void issue_probe(int device, double* gops, double* clock_mhz);
void launch_grand_product(const uint32_t* const* num, const uint32_t* const* den, uint32_t* const* ratio,
                          uint32_t* const* gp, int n, int batch, uint32_t* scan_tmp, hipStream_t s);

// Lookups and logup (board/module.go:56-71, poly/lookup_multiplicity.go:37-100,
// board/logup.go:138-186), batch on grid y, per-instance pointer arrays.
// (The range table column itself is a generated column: launch_gencol.)
void launch_fill(uint32_t* const* p, int n, uint32_t v, int batch, hipStream_t s);
// counts[b][v] += 1 for every canonical value v < bound of src[b][0..n), *err[b] += 1 for the others
// (several batch entries may share a count table and an error word: atomic additions)
void launch_histogram(const uint32_t* const* src, int n, uint32_t* const* counts, uint32_t bound,
                      uint32_t* const* err, int batch, hipStream_t s);
void launch_counts_to_field(uint32_t* const* c, int n, int batch, hipStream_t s);

// General lookup multiplicities (poly/lookup_multiplicity.go:37-100): tuples
// of any number of base-field columns with optional selectors, several
// sources and targets.  Per instance, a hash table keyed by tuple, whose
// slots hold a representative target row (rows are compared word by word
// through the tables, so a slot is claimed by one compare-and-swap): the
// first selected target row of each tuple (targets in order, then rows),
// the selected source weight of each tuple (a 64-bit sum of canonical
// selector values, or of ones), and the multiplicity weight / selT written
// at the first row.  Source tuples found in no target are summed in a second
// table (with source selectors) or flag the instance's error word at once.
struct LkTable {  // one source or target table of one instance
  const uint32_t* const* col;     // w tuple columns (Montgomery, base field), a device array
  const uint32_t* sel;            // selector column, or null (all ones)
  uint32_t* mult;                 // targets: the multiplicity column
  uint32_t n, off;                // rows; first global row index (targets, then sources, counted apart)
};
struct LkJob {  // one instance
  uint32_t* slot_row;             // cap: representative target row + 1, 0 empty
  uint32_t* slot_first;           // cap: first selected target row (init ~0)
  unsigned long long* slot_w;     // cap: summed source weight (init 0)
  uint32_t* miss_row;             // miss_cap: representative source row + 1 (null: no selectors)
  unsigned long long* miss_w;     // miss_cap
  uint32_t* err;                  // error word
  uint32_t cap_log, miss_cap_log;
};
// tabs: per instance nt targets then ns sources (batch x (nt + ns)); counts
// of target and source rows per instance: total_t, total_s
void launch_lookup_general(const LkTable* tabs, const LkJob* jobs, int nt, int ns, int w, uint32_t total_t,
                           uint32_t total_s, uint32_t max_cap, uint32_t max_miss_cap, int batch, hipStream_t s);
// acc = (first ? 0 : acc) + m / e (E6 columns, limb stride n)
void launch_logup_acc(const uint32_t* const* e, const uint32_t* const* m, uint32_t* const* acc, int n, int first,
                      int batch, hipStream_t s);
// col = inclusive running sum of acc minus (i+1) * total / n; total to tot[6b..6b+6)
void launch_cyclic_logup_finish(const uint32_t* const* acc, uint32_t* const* col, uint32_t* tot, int n,
                                uint32_t ninv, int batch, uint32_t* scan_tmp, hipStream_t s);

// Quotient coefficients from the per-coset inverse transforms:
//   coef[b][t][l][a] = ginv[a + t n] * sum_i wr[(t i) mod rho] * Y[b][l][i n + a]
// (wr[k] = w_rho^-k, ginv[m] = g^-m).
void launch_quot_combine(const uint32_t* Y, uint32_t* coef, const uint32_t* ginv, const uint32_t* wr, int rho,
                         int n, int batch, hipStream_t s);

// Lagrange weights at zeta: W[i] = common * w^i / (zeta - w^i), E6 as 6 limb
// arrays (stride n).  zeta, common: 6 words per instance.
void launch_lagrange_weights(const uint32_t* zeta, const uint32_t* common, const uint32_t* wpow,
                             uint32_t* const* W, int n, int batch, hipStream_t s);

// Claimed value jobs: value = sum_i W[(i - shift) mod n] * P[i].
// Claimed values (internal/fri/values.go:46-110): value = sum_j W[j] P[(j + s) mod n]
// with W the Lagrange weights L_j(zeta) of the polynomial's module.  The
// claims of one (instance, module) share W: a ClaimSet lists them, so W is
// read once per row for up to CLAIMS_GROUP claims at a time.
struct ClaimItem {
  const uint32_t* poly;  // 1 (base) or 6 (ext) limb arrays, stride n
  uint32_t* out;         // 6 words
  int32_t shift;
  int32_t ext;
};
struct ClaimSet {
  const uint32_t* W;     // 6 limb arrays, stride n
  int32_t first, count;  // items [first, first + count)
};
constexpr int CLAIMS_GROUP = 8;
// Rows of a set are split over claims_split(n) blocks; part then holds
// n_items x claims_split(n) x 6 words of partial sums (item-major).
constexpr int CLAIMS_MAX_SPLIT = 4096;
int claims_split(int n, int rows_per_block = 1024);
void launch_claims(const ClaimSet* sets, int n_sets, const ClaimItem* items, int n_items, int n, uint32_t* part,
                   hipStream_t s, int rows_per_block = 1024);

// Launch choices of a prove that do not change its result, picked per
// (plan, batch) by Plan::autotune: proving lanes, the width x batch from
// which Merkle levels run one launch per level (below it, cooperative fused
// launches), threads per block of the leaf and level hashing kernels, the FRI
// layer size (leaf pairs) up to which the fold and leaf hashing run
// cooperatively, and rows per block of the claimed values.
struct Tuning {
  int lanes = 2;
  int merkle_wide_min = 1 << 15;
  int hash_threads = 256;
  int fri_coop_pairs = 512;
  int claims_rows = 1024;
};

// DEEP quotient on the trace domain for one class: per row x (natural order)
//   DQ[x] = sum_p (c_p - sum_{bundles i at p} a_i * P_i(x)) / (z_p - w^x)
struct DeepBundle {
  int32_t ext;          // rail of P_i
  int32_t col;          // column index into the per-instance column table
  int32_t first_pt;     // index into pt_list
  int32_t n_pts;
};
struct DeepDesc {
  int n, n_bundles, n_points;
  const DeepBundle* bundles;
  const int32_t* pt_list;     // point indexes used by the bundles
  const uint32_t* const* cols;  // [b * n_cols + col]
  int n_cols;
  const uint32_t* scale;      // [b][bundle][6]  alpha^i
  const uint32_t* consts;     // [b][point][6]   c_p
  const uint32_t* points;     // [b % points_mod][point][6]   z_p
  int points_mod;             // 0: points per launch instance b
  const uint32_t* wpow;       // w^x
  uint32_t* const* out;       // [b] 6 limb arrays, stride n
};
void launch_deep(const DeepDesc& d, int batch, hipStream_t s);
// alpha_DEEP derived on the device with the DEEP constants of a single class
// (see k_deep_alpha): prev: the previous challenge's digest per instance;
// cvals: [b][ncv][6] claimed values, order: binding order (ncv indexes);
// terms: (bundle, claimed value, point) triples of the class.
struct DeepAlphaArgs {
  int n_pre;
  const uint32_t* prev;
  const uint32_t* cvals;
  const int32_t* order;
  int ncv;
  const int3* terms;
  int n_terms, n_bundles, n_points;
  uint32_t* dig;     // [b][8] the alpha_DEEP digest
  uint32_t* scale;   // [b][bundle][6]
  uint32_t* consts;  // [b][point][6]
};
void launch_deep_alpha(const DeepAlphaArgs& a, const uint32_t* prefix, int n_prefix, int batch, hipStream_t s);

// Generated columns (module.go:77-101, 229-293) in one launch, one job per
// column: GENCOL_RANGE 0..bound-1 then zeros, GENCOL_LAGRANGE the unit vector
// at a0, GENCOL_ID g^a0 w^j, GENCOL_SIGMA g^(t >> logn) w^(t mod n) for
// t = S[b][a0 n + j].  Job k covers blocks [block0, block0 + ceil(n/256)) in
// x; y is the instance.
enum GenColKind : int { GENCOL_RANGE = 0, GENCOL_LAGRANGE = 1, GENCOL_ID = 2, GENCOL_SIGMA = 3 };
struct GenColJob {
  int kind, n, logn;
  uint32_t a0;               // range bound, Lagrange position or sigma chunk
  uint32_t gk;               // identity support: g^a0
  const uint32_t* wpow;      // identity support, sigma: powers of w_n
  const uint32_t* gpow;      // sigma: powers of g
  uint32_t* const* outs;     // B column blocks
  const uint32_t* const* S;  // sigma: B permutations
  uint32_t len;              // sigma: permutation length (entries must be below it)
  uint32_t* err;             // sigma: B flags, set to 1 for an instance with an entry >= len
  int block0;
};
void launch_gencol(const GenColJob* jobs_dev, int njobs, int total_blocks, int batch, hipStream_t s);
// dst[b][rows[i]] = src[b][i], i < k (rows checked against the column size by
// the plan parser).
void launch_scatter_rows(const uint32_t* const* src, const uint32_t* rows, int k, uint32_t* const* dst, int batch,
                         hipStream_t s);

// Fiat-Shamir challenges on the device (transcript.go:285-315): the sponge
// of a message given as segments; word i of segment k for instance b is
// seg[k].p[b * seg[k].stride + i] (stride 0: the same words for every
// instance).  Writes dig[8 b ..] and, if e6, the first six words to
// e6[b * e6_stride ..].
struct FsSeg {
  const uint32_t* p;
  uint32_t len;
  uint32_t stride;
};
// seg_dev: nseg segments in device memory; len: their total length.
void launch_fs_challenge(const FsSeg* seg_dev, int nseg, int len, uint32_t* dig, uint32_t* e6, int e6_stride,
                         int batch, hipStream_t s);
// "FSPW": the tag before a proof of work's bit count and salt
// (internal/fiat-shamir/transcript.go:36)
constexpr uint32_t TAG_FSPW_WORD = 0x46535057u;
// Proof-of-work salt search of one challenge per instance: the message is
// the segments followed by [FSPW, bits, salt]; found[b] (set to ~0 by the
// caller) receives the smallest salt below p whose digest has `bits` zero
// low bits (~0 if none), and trailer[3 b ..] the message's last three words
// for that salt (Montgomery).  `workgroups` workgroups of 256 salts each
// search one instance.
void launch_fs_grind(const FsSeg* seg_dev, int nseg, int len, uint32_t bits, uint32_t* found, uint32_t* trailer,
                     int workgroups, int batch, hipStream_t s);
// out[b][k] = chal[b][ids[k]] for the nch E6 challenge operands of a DAG.
void launch_chal_compact(const uint32_t* chal, int ncols, const uint32_t* ids, int nch, uint32_t* out, int batch,
                         hipStream_t s);
// chal[b][col] = tot[b] * s_mont (E6 times a base-field element).
void launch_scale_into(const uint32_t* tot, uint32_t s_mont, uint32_t* chal, int ncols, int col, int batch,
                       hipStream_t s);
// z[b] = zeta_b and cm[j][b] = (zeta_b^sizes[j] - 1) * ninv[j], zeta_b the
// first six words of zdig[8 b ..].
void launch_zeta_consts(const uint32_t* zdig, const uint32_t* sizes, const uint32_t* ninv, int nmod, uint32_t* z,
                        uint32_t* cm, int batch, hipStream_t s);
// out[b][p] = zeta_b * w[p].
void launch_zeta_points(const uint32_t* zdig, const uint32_t* w, int np, uint32_t* out, int batch, hipStream_t s);

// FRI fold (fri.go:1177-1198): next[i] = (p+q)/2 + alpha*(p-q)/2*xinv[i],
// p = layer[2i], q = layer[2i+1]; E6 layers as 6 limb arrays.
// FRI fold challenge per instance on the device: digest of the sponge over
// prefix (host memory, n_prefix <= 32 words) || prev[8 b ..] || root of
// trees[b]; writes dig[8 b ..], alpha[6 b ..] (first 6 limbs) and the root
// to roots[8 b ..].
void launch_fri_challenge(const uint32_t* prefix, int n_prefix, const uint32_t* prev, uint32_t* const* trees,
                          uint32_t* dig, uint32_t* alpha, uint32_t* roots, int batch, hipStream_t s);
// FRI fold into out (length half) fused with the leaf hashes of out's tree
// (row pairs, one E6 column) written to leaves[b] (half/2 digests).
// Layers of at most coop_pairs leaf pairs use the cooperative form.
void launch_fri_fold_leaf(const uint32_t* const* in, uint32_t* const* out, const uint32_t* alpha,
                          const uint32_t* xinv2, int half, uint32_t hdr2, uint32_t* const* leaves, int batch,
                          hipStream_t s, int coop_pairs = 512);
void launch_fri_fold(const uint32_t* const* in, uint32_t* const* out, const uint32_t* alpha, const uint32_t* xinv2,
                     int half, int batch, hipStream_t s);
// running[k] += gamma * level[k]
void launch_fri_combine(uint32_t* const* running, const uint32_t* const* level, const uint32_t* gamma, int len,
                        int batch, hipStream_t s);

// Query openings of a range of instances in one launch.  Each op copies, for
// every (instance b, query q), a fixed number of units into the record of
// (b, q) at out + (b * nq + q) * stride + out_off; blocks [block0, block0 +
// nblocks) of the grid belong to the op.
//   rows:   units = 2 * ncols words: column c at rows lo, lo + 1 (lo first),
//           row = pos >> shift, lo = row & ~1, columns ptrs[b * ncols + c]
//   path:   units = ncols (= depth) digests: siblings of leaf lo / 2 of the
//           tree ptrs[b] with nleaves leaves, leaf level first
//   digest: units = 1 digest: ptrs[b] + 8 * ((lo / 2) >> d)
struct GatherOp {
  int32_t kind;  // 0 rows, 1 path, 2 digest
  int32_t shift;
  int32_t ncols;
  int32_t d;
  int32_t nleaves;
  int32_t out_off;
  int32_t block0;
  int32_t nblocks;
  const uint32_t* const* ptrs;
};
// Instances b0 .. b0 + nb - 1 of the batch; each op's blocks cover at least
// nb * nq records.
void launch_gather(const GatherOp* ops_dev, int n_ops, int total_blocks, const uint32_t* pos, int nq, int b0, int nb,
                   uint32_t* out, int stride, hipStream_t s);

// Query gathers.  pos: [b * nq + q] full-domain positions.
// Rows: row = pos >> shift, lo = row & ~1; writes cols[c][lo], cols[c][lo+1]
// for every column c to out + (b * nq + q) * stride + off (lo values then hi).
void launch_gather_rows(const uint32_t* pos, int nq, int shift, const uint32_t* const* cols, int ncols,
                        uint32_t* out, int stride, int off, int batch, hipStream_t s);
// Merkle path of leaf ((pos >> shift) & ~1) / 2: depth siblings (8 words each).
void launch_gather_path(const uint32_t* pos, int nq, int shift, const uint32_t* const* trees, int nleaves,
                        uint32_t* out, int stride, int off, int batch, hipStream_t s);
// Injection leaves: digest inj[b][((pos >> shift) & ~1) / 2 >> d].
void launch_gather_digest(const uint32_t* pos, int nq, int shift, int d, const uint32_t* const* inj,
                          uint32_t* out, int stride, int off, int batch, hipStream_t s);

}  // namespace lg
