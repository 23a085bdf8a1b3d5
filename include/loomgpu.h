/*
 * loomgpu -- C ABI of the MI355X prove phase for loom (KoalaBear / Poseidon2 /
 * multi-degree FRI).
 *
 * The reference has no FFI layer: its prove path sits behind the Go function
 *   prover.Prove(t trace.Trace, pk setup.ProvingKey, pi public.Inputs,
 *                program board.Program, opts ...prover.Option) (proof.Proof, error)
 * (prover/prover.go:655-719; loom.Prove wraps it, statement.go:91-97).  The entry
 * points below are what a cgo shim of that function binds: the shim
 * serialises the compiled program (board.Compile, program.go:45-89) and its
 * commitment layout (protocol.BuildLayout / BuildCanonicalSchedule,
 * internal/protocol/layout.go:165-241, schedule.go:92-235) into a plan once,
 * then hands trace columns (Montgomery u32, the in-memory form of
 * []koalabear.Element) to lg_prove and decodes the canonical proof words back
 * into proof.Proof (proof/proof.go:31-51).
 *
 * Conventions: every function returns 0 on success and a negative code on
 * failure; lg_last_error() returns the failure message of the calling thread.
 * No entry point retains a caller pointer after it returns.  A context owns
 * one HIP stream on one device; calls on one context are serialised by the
 * library, several contexts may run concurrently.
 *
 * What of prover.Prove (prover/prover.go:55-108, 655-719) this boundary
 * carries, and what it refuses:
 *   - trace.Trace: the input columns (base field) and the copy-constraint
 *     permutations; the prover steps' columns are computed on the device.
 *   - board.Program + layout: the plan word stream.
 *   - WithNumFriQueries: the plan's num_queries.
 *   - WithHashBackend: Poseidon2 only (hash id "poseidon2").
 *   - WithPhaseCallback: lg_plan_last_phases after the call.
 *   - SkipFRI: lg_plan_set_options(plan, LG_PROVE_SKIP_FRI).
 *   - WithFriOption(fri.WithGrinding(bits)): lg_plan_set_grinding.
 *   - setup.ProvingKey: lg_plan_set_setup commits the setup columns once per
 *     plan and keeps them resident (expr.Setup and expr.ExtSetup columns).
 *   - public.Inputs: entries at rows the plan lists, base-field values or
 *     (for inputs declared with expr.PublicInputExt) E6 values; each
 *     instance passes the values after its input columns.
 *   - WithTranscript: a fresh running transcript is lg_prove (its computed
 *     challenges come back in `transcripts`); one with computed challenges is
 *     lg_prove_chained (its last digest).
 *   - Refused (a plan or call that needs them is rejected with an error):
 *     extension entries of a public input declared in the base field,
 *     WithNewTranscriptHasher (the transcript hasher is the Poseidon2 sponge),
 *     other hash backends.  (fri.WoFullDomainAllocation, the other FRI
 *     option, changes only the reference's memory use: nothing to carry.)
 *   - Lookups: every arguments/lookup.go form (Range, Lookup, LookupTuple,
 *     LookupUnion(Tuple), CLookup(Tuple), CLookupUnion(Tuple)): tuples of up
 *     to 8 base-field expressions, optional selectors, targets in several
 *     modules; a source tuple in no selected target fails its instance.
 *
 * Plan word stream: see loom_amd/plan.py (module sizes, column table, rounds
 * and prover steps, commitment trees with opening shifts, folded constraint
 * DAGs).  lg_plan_create and lg_plan_check check every index in it.
 *
 * Proof word stream (all u32; field elements Montgomery; E6 = 6 words in the
 * order B0.A0 B0.A1 B1.A0 B1.A1 B2.A0 B2.A1).  It encodes proof.Proof
 * (proof/proof.go:31-51) with every map in sorted-key order:
 *   hash id (byte length, then bytes packed 4 per word, little endian)
 *   #exposed, then per exposed value in ascending byte order of its name
 *     (proof.ExposedValues, proof/public_inputs.go:23-49):
 *       name (byte length, then bytes packed 4 per word, little endian)
 *       #entries, then per entry: Idx, ExtValue() as E6 (6 words; a base
 *       entry lifted to E6)
 *     (the logup totals of cyclic logup columns: one entry, Idx 0, each)
 *   #commitments, roots (8 words each; trace rounds then AIR)
 *   claimed values (fri.OpeningProof.ClaimedValues):
 *     #trees { #groups { #base { #shifts, E6 per shift } #ext { #shifts, E6 per shift } } }
 *   FRI proof (fri.Proof):
 *     levels root (8 words)
 *     #levels - 1, then per level l >= 1: #queries { opening }   (no path)
 *     #roots, roots (8 words each; one per folded layer)
 *     1 (the final polynomial is extension-field), #final, E6 per value
 *     #queries, then per query: #layers { opening }
 *     #proofs of work (0 without grinding), then per fold challenge in
 *       ascending byte order of its name ("fri_fold_0", "fri_fold_1",
 *       "fri_fold_10", ...): name (byte length, then bytes packed 4 per
 *       word), NbBits, Salt (Montgomery)
 *   samplings (PointSamplings): #queries { #trees { row(lo), row(hi), path,
 *     #injected groups { row(lo), row(hi) } } }
 *   opening := 1 (extension-field layer), row index, P (E6), Q (E6), path:
 *     the layer's values at rows 2k and 2k+1 (P at the even row)
 *   path := leaf index, #siblings, siblings (8 words each, leaf to root),
 *           #injection leaves, leaves (8 words each)
 *   row  := #base, base values, #ext, ext values (6 words each)
 */
#ifndef LOOMGPU_H
#define LOOMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lg_ctx lg_ctx;
typedef struct lg_plan lg_plan;

/* Message of the last failed call on this thread ("" if none). */
const char* lg_last_error(void);

/* Context on HIP device `device` (one stream at default priority). */
int lg_ctx_create(int device, lg_ctx** out);
void lg_ctx_destroy(lg_ctx* ctx);

/* Parses a plan and compiles its constraint DAGs (hiprtc, gfx950).  The
 * compile step replaces the reference's per-row DAG interpreter
 * (internal/dag/dag.go:1134-1446) and runs once per distinct DAG. */
int lg_plan_create(lg_ctx* ctx, const uint32_t* words, size_t n_words, lg_plan** out);
void lg_plan_destroy(lg_plan* plan);
/* Parses a plan and compiles its DAG kernels with hiprtc without touching a
 * device (no context needed): a validation step a caller can run at
 * board.Compile time.  Returns the total code-object bytes in *code_bytes. */
int lg_plan_check(const uint32_t* words, size_t n_words, size_t* code_bytes);

/* Number of input arrays / permutations an instance must provide: the trace
 * input columns (N words each, plan input order), then for each public-input
 * column of the plan (public.Inputs, name order) its entry values in
 * ascending row order: one word per base entry, six (the E6 limbs B0.A0 ..
 * B2.A1, Montgomery) per extension entry, as the plan marks each entry. */
int lg_plan_num_inputs(const lg_plan* plan, size_t* n_columns, size_t* n_perms);

/* setup.Setup (setup/setup.go:68-124) for a plan with setup columns
 * (expr.Setup leaves, e.g. PLONK selectors, plonk_gnark_utils.go:217-219):
 * setup_columns[k] is setup column k in plan setup order (host memory,
 * Montgomery): N words of its module, or for an extension column
 * (expr.ExtSetup) its six limb arrays of N words one after another (limb
 * order B0.A0 .. B2.A1).  The columns are RS-encoded and committed into
 * the setup tree (one group per module) once; columns, LDE and tree stay in
 * device memory for every later prove of the plan (the proving key), which
 * binds the root to its first challenge and opens the tree at the queries.
 * root: 8 words, the verification key (setup.VerificationKey.Roots[0]).
 * Calling it again replaces the key.  A plan with setup columns refuses to
 * prove before its key is set. */
int lg_plan_set_setup(lg_plan* plan, const uint32_t* const* setup_columns, size_t n_columns, uint32_t* root);

/* Proves `batch` independent instances of one plan (same shapes, different
 * witnesses), in lockstep: one launch sequence per phase for all instances.
 * With batch = 1 this is prover.Prove for one statement.
 * An instance with an invalid witness (a lookup value outside its table, a
 * permutation entry outside its cells) fails alone: its proof_lens[b] is 0,
 * the other instances' proofs are written, and the call returns -3 with
 * lg_last_error() naming each failed instance ("instance b: ...").
 *   columns[b * n_columns + k]: host pointer, N words (Montgomery), column k of
 *                               instance b in plan input order
 *   perms[b * n_perms + p]:     host pointer, permutation p of instance b
 *                               (the board.PermutationGen S slice, as u32)
 *   proofs[b], proof_caps[b]:   caller buffers for the proof words
 *   proof_lens[b]:              receives the proof length in words
 *   transcripts[b]:             optional (may be NULL): receives the challenge
 *                               digests (8 words each) in computation order,
 *                               capacity transcript_caps[b] words
 * Returns -2 when a buffer is too small (lengths are still reported), -3 when
 * some instances failed (see above), -1 on any other failure. */
int lg_prove(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
             uint32_t* const* proofs, const size_t* proof_caps, size_t* proof_lens, uint32_t* const* transcripts,
             const size_t* transcript_caps, size_t* transcript_lens);

/* lg_prove with trace inputs already in HBM: columns[] and perms[] are device
 * pointers on the context's device (checked with hipPointerGetAttributes),
 * read during the call only.  Same proofs as lg_prove; the caller that keeps
 * its trace device-resident skips the host->device copy of the inputs. */
int lg_prove_device(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                    uint32_t* const* proofs, const size_t* proof_caps, size_t* proof_lens,
                    uint32_t* const* transcripts, const size_t* transcript_caps, size_t* transcript_lens);

/* prover.WithTranscript(fs) (prover/prover.go:54-60, 196-205) with a running
 * transcript whose last challenge is computed: the prove's challenges are
 * registered after the caller's, so its first challenge's message carries
 * that challenge's digest after its name, as every later one carries its
 * predecessor's (internal/fiat-shamir/transcript.go:285-315).  fs_prev:
 * batch x 8 words (Montgomery), instance b's digest at 8 b.  device_inputs: 0
 * as lg_prove, 1 as lg_prove_device.  (A fresh running transcript is lg_prove
 * itself; the challenges a prove computed come back in `transcripts`.) */
int lg_prove_chained(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                     int device_inputs, const uint32_t* fs_prev, uint32_t* const* proofs, const size_t* proof_caps,
                     size_t* proof_lens, uint32_t* const* transcripts, const size_t* transcript_caps,
                     size_t* transcript_lens);

/* Autotuner (no counterpart in the reference).  Launch choices that do not
 * change the proofs -- proving lanes, the Merkle width from which levels run
 * one launch per level, threads per block of leaf and level hashing, the FRI
 * layer size up to which fold and leaf hashing run cooperatively, rows per
 * block of the claimed values -- are picked per (plan, batch) by timing whole
 * proves of the given instances (inputs as for lg_prove, or device pointers
 * with device_inputs = 1) over a coordinate descent, `reps` timed proves per
 * trial (median).  The choice is kept by the plan for proves of that batch
 * size; with LOOMGPU_TUNE_FILE set it is appended to that table, which
 * lg_plan_create reads back for the same plan and device.
 *   tuning_out: 6 ints: lanes, merkle_wide_min, hash_threads, fri_coop_pairs,
 *               claims_rows, trials;  ms_out: 2 doubles: the starting and the
 *               chosen configuration's median prove time. */
int lg_plan_autotune(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                     int device_inputs, int reps, int32_t* tuning_out, double* ms_out);
/* Prove options for the proves of this plan that follow (flags):
 *   LG_PROVE_SKIP_FRI  prover.SkipFRI (prover/prover.go:78-83, 626-653,
 *     691-700): the proof carries the exposed values, commitments and claimed
 *     values; the rest of the opening is empty (levels root 0, no levels, no
 *     roots, an empty final polynomial "1, 0", no queries, no proofs of work,
 *     no samplings).  lg_verify with skip_fri checks such proofs.
 * Other flags are refused. */
#define LG_PROVE_SKIP_FRI 1u
int lg_plan_set_options(lg_plan* plan, uint32_t flags);
/* prover.WithFriOption(fri.WithGrinding(bits)) (prover/prover.go:62-68,
 * internal/fri/fri.go:71-78, 1053-1090): every FRI fold challenge of later
 * proves (and the challenges lg_verify replays) carries a proof of work of
 * `bits` bits (0: off, the default; at most 31).  The salt search runs on the
 * device; the proof lists one (name, bits, salt) per fold challenge. */
int lg_plan_set_grinding(lg_plan* plan, int bits);
/* Sets the launch choices for proves of `batch` instances (5 ints, as above). */
int lg_plan_set_tuning(lg_plan* plan, int batch, const int32_t* tuning);

/* HIP graphs per shape bucket (on by default; LOOMGPU_GRAPHS=0 turns the
 * default off).  For a plan whose modules share one size (one DEEP class), the
 * device phase of a prove (generated columns through the FRI commit phase)
 * needs no host round trip: the second prove of a shape (batch size) records
 * it into a HIP graph over buffers kept for that shape, later proves of the
 * shape replay the graph.  Proofs are the same either way.  At most 4 shapes
 * per plan are kept (least recently used dropped). */
int lg_ctx_set_graphs(lg_ctx* ctx, int on);
/* Graph use of a plan so far (4 counts): device bytes held by its shape
 * buckets, proves that recorded a graph, proves that replayed one, and proves
 * that ran again eagerly because a recording or a replay failed (a replay
 * fails when a table it would stage differs from the recorded one). */
int lg_plan_graph_stats(const lg_plan* plan, uint64_t* out4);

/* Per-kernel device timing with HIP events on the context's stream (off by
 * default).  lg_ctx_timing_report writes one text line per kernel name,
 * "name launches device_ms algorithmic_bytes poseidon2_permutations", into
 * buf (NUL-terminated, *len = bytes needed) and clears the totals. */
int lg_ctx_set_timing(lg_ctx* ctx, int on);
int lg_ctx_timing_report(lg_ctx* ctx, char* buf, size_t cap, size_t* len);

/* Phase timings of the last lg_prove on this plan (seconds): execute-steps,
 * compute-air-quotients, pcs-open (pcs-claimed-values-only under SkipFRI):
 * the phases prover.WithPhaseCallback reports, prover.go:682-716. */
int lg_plan_last_phases(const lg_plan* plan, double* out3);

/* ---- serving ---- */

/* A serving loop on one device (no counterpart in the reference, whose
 * bench proves sequentially; it is the request loop a prover service puts in
 * front of prover.Prove).  `slots` worker threads (1..16) each own a context
 * and stream.  A request is admitted, oldest first, when a slot is free and
 * its HBM footprint (lg_plan_bytes_per_instance x batch) fits within
 * hbm_budget_bytes beside the running requests (0: 80 % of the free device
 * memory at creation); a request larger than the budget runs alone.  Queued
 * requests of the same plan join the prove of the one admitted, up to
 * max_batch instances (continuous batching); every proof equals the one
 * lg_prove gives for that instance alone.  Failures are per request: a request
 * with an invalid instance fails (lg_server_wait returns -1 naming it) while
 * the requests merged into the same prove complete normally. */
typedef struct lg_server lg_server;
int lg_server_create(int device, int slots, uint64_t hbm_budget_bytes, int max_batch, lg_server** out);
/* Stops the workers after their current prove; queued requests fail. */
void lg_server_destroy(lg_server* server);
/* Queues `batch` instances of the plan (inputs as for lg_prove, host memory,
 * copied before return); *ticket identifies the request. */
int lg_server_submit(lg_server* server, const uint32_t* plan_words, size_t n_words, int batch,
                     const uint32_t* const* columns, const uint32_t* const* perms, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request.  Returns 0 with the
 * proofs copied out (the ticket is then released), 1 on timeout, -2 if a
 * buffer is too small (lengths reported, the result is kept for another
 * call), -1 if the request failed (message in lg_last_error). */
int lg_server_wait(lg_server* server, uint64_t ticket, int timeout_ms, uint32_t* const* proofs,
                   const size_t* proof_caps, size_t* proof_lens);
/* out[0..n): submitted, completed, failed, queued, running, peak running,
 * peak admitted bytes, proves launched, requests merged into another's
 * prove, budget bytes. */
int lg_server_stats(lg_server* server, uint64_t* out, size_t n);
/* Admission footprint estimate of one instance of a plan (no device). */
int lg_plan_bytes_per_instance(const uint32_t* words, size_t n_words, uint64_t* bytes);

/* ---- building blocks (device pointers), the minimum slice of the path ---- */

/* Poseidon2 row-pair leaf hashing of one committed group
 * (Poseidon2LeafHasher.HashLeafPair, internal/fri/hash.go:39-56): for pair
 * leaf k < rows/2, hashes [LEAF, 2nb, 2ne, lo base, hi base, lo ext, hi ext]
 * where base column c is base_cols[c] and ext column e is the 6 limb arrays
 * ext_cols[e] + l*rows (l < 6), rows 2k and 2k+1.  digests: rows/2 x 8 words. */
int lg_hash_leaf_pairs(lg_ctx* ctx, const uint32_t* const* base_cols, int nb, const uint32_t* const* ext_cols, int ne,
                       size_t rows, uint32_t* digests);

/* Merkle tree over n_leaves leaf digests (merkle.Tree.Build,
 * internal/merkle/tree.go:182-296, Poseidon2NodeCompress nodes): nodes is the
 * 1-indexed flat array of 2*n_leaves digests whose leaves (nodes[n_leaves..])
 * are already filled; on return nodes[1] is the root.  Level injections:
 * inj[k] (inj_width[k] digests, non-increasing widths) fold into the level of
 * that width (merkle.LevelInjection). */
int lg_merkle_build(lg_ctx* ctx, uint32_t* nodes, size_t n_leaves, const uint32_t* const* inj, const size_t* inj_width,
                    int n_inj);

/* Grand product (board.GrandProductStep, poly/iop_utils.go:170-192, 315-322):
 * gp[0] = 1, gp[i] = gp[i-1] * num[i-1] / den[i-1] over E6 columns given as
 * 6 limb arrays of n words each (limb l at +l*n), n a power of two. */
int lg_grand_product(lg_ctx* ctx, const uint32_t* num, const uint32_t* den, uint32_t* gp, size_t n);

/* lg_grand_product without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same inputs and result as lg_grand_product (device pointers);
 * the ratio column and the scan partials live in a file-scope device array of
 * about 6 MiB per device, held until the process ends, so n is limited to
 * LG_NOALLOC_MAX_ROWS rows (a larger n is refused with an error).  The array
 * is shared by every call on one device: calls that overlap on one device
 * (from different contexts or threads) race; calls on one context are
 * serialised by the library. */
#define LG_NOALLOC_MAX_ROWS (1u << 18)
int lg_grand_product_noalloc(lg_ctx* ctx, const uint32_t* num, const uint32_t* den, uint32_t* gp, size_t n);

/* Diagnostic (no counterpart in the reference): the device's issue rate of
 * the integer instructions the field arithmetic uses, measured with eight
 * independent chains per lane.  out5[0..3]: lane operations per ns of
 * v_add_u32, v_mul_lo_u32, v_mad_u64_u32 and the Montgomery product;
 * out5[4]: the device clock in MHz.  The VALU roofline of the Poseidon2
 * kernels is priced against these rates. */
int lg_probe_issue_rate(lg_ctx* ctx, double* out5);

/* RS encoding (reedsolomon.Encoder.Encode, internal/reedsolomon/rs.go:52-70) of
 * n_cols columns of n words each (Lagrange form on the size-n subgroup, one
 * column after the other) to 4n bit-reversed evaluations each. */
int lg_rs_encode(lg_ctx* ctx, const uint32_t* in, size_t n, int n_cols, uint32_t* out);

/* verifier.Verify (verifier/verifier.go:479-533) for n proofs of this plan.
 * The host decodes each proof, replays its transcript and checks the logup
 * buses and every module's relation at zeta; the device hashes every opened
 * row and walks every Merkle path of the batch (committed trees with their
 * level injections, the FRI levels tree and layers) while the host checks the
 * DEEP bridge and the FRI folds (internal/fri/verify.go:60-157,
 * internal/fri/fri.go:751-1534).
 *   proofs[i], lens[i]: proof words (the grammar above);
 *   setup_roots: the verification key, 8 Montgomery words per setup tree (the
 *     root lg_plan_set_setup returned), n_setup_roots of them (0 and NULL
 *     without setup columns);
 *   public_values: NULL without public inputs, else n * n_public pointers:
 *     [i * n_public + j] -> the entry values of public input j of proof i
 *     (Montgomery, as lg_prove takes them);
 *   skip_fri: nonzero stops after the AIR check (verifier.SkipFRI, :52-57);
 *   status[i]: 0 valid, else the failed check: 1 proof format or shape,
 *     2 logup bus, 3 AIR identity at zeta, 4 opening of a committed tree,
 *     5 DEEP bridge, 6 FRI (layer opening, fold or final polynomial).
 * Returns 0 once every proof has its verdict, whatever the verdicts. */
int lg_verify(lg_plan* plan, int n, const uint32_t* const* proofs, const size_t* lens, const uint32_t* setup_roots,
              size_t n_setup_roots, const uint32_t* const* public_values, int skip_fri, int32_t* status);
/* verifier.WithTranscript (verifier/verifier.go:60-65, 167-176) for proofs made
 * with lg_prove_chained: fs_prev as there, n x 8 words. */
int lg_verify_chained(lg_plan* plan, int n, const uint32_t* const* proofs, const size_t* lens,
                      const uint32_t* setup_roots, size_t n_setup_roots, const uint32_t* const* public_values,
                      int skip_fri, const uint32_t* fs_prev, int32_t* status);
/* The message of proof i's failed check in this plan's last lg_verify ("" when
 * it passed); NUL-terminated into buf, *len = bytes needed. */
int lg_plan_verify_message(const lg_plan* plan, int i, char* buf, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif

#endif /* LOOMGPU_H */
