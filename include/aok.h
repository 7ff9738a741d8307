/*
 * aok.h -- C ABI of libaok.so, the MI355X (gfx950) kernels behind the two hot paths
 * of AtomicOrbitalKernels.jl:
 *   H1  batched atomic-orbital evaluation  P[i,o] = R_{iR[o]}(r_i; sigma_i) Z_{iY[o]}(x_i),
 *       its spatial gradients and its (zeta, D, x) pullback;
 *   H2  batched 2-centre and 3-centre Cartesian-Gaussian overlaps (McMurchie-Davidson)
 *       and the 2-centre (zeta, D) pullback.
 *
 * Conventions (all entry points)
 *  - Bulk arrays (positions, species indices, parameters, outputs, cotangents) are
 *    DEVICE pointers, in the reference's Julia column-major layouts:
 *        X, dX          [3*nX]             xyz interleaved (Vector{SVector{3,T}})
 *        zeta, D, coef  [n*K*NZ]           element (k, m, sigma) at k + n*(m + K*sigma)
 *        P, dPin        [nX*nOrb]          element (i, o) at i + nX*o
 *        dP             [3*nX*nOrb]        element (i, o, c) at 3*(i + nX*o) + c
 *        posA/B/C       [3*B]
 *        out (2c)       [N*N*B]            each out[:,:,b] contiguous
 *        out (3c)       [N*N*N*B]
 *    Small structural integer arrays (poly, iR, iY, ls) are HOST pointers (int64, the
 *    reference keeps them as host Vector{Int}); iR/iY are 1-based like Julia's.
 *  - Species indices sidx are 1-based; itype selects int64 (Julia Int, zero-copy)
 *    or int32.  A NULL sidx means "every point is species 1" (the reference's
 *    default for plain-coordinate input, gtostoradials.jl:71-72).
 *  - Floating dtype: AOK_F32 or AOK_F64 for every float array of a call.
 *  - Ownership: the caller allocates every output; the library never retains a
 *    caller pointer past the call.  Calls are blocking (the reference synchronises
 *    after every wrapper): results are complete when the call returns.
 *  - Errors: every entry point returns an aok_status and never throws across the ABI;
 *    aok_last_error(ctx) gives the message.  Validation happens before any launch,
 *    except the per-element species range check, which the kernels perform and the
 *    call reports (AOK_ERR_SPECIES) after its closing synchronisation.
 *  - A context owns one HIP stream (or borrows the caller's), its workspace and its
 *    autotune table.  An owned stream is a blocking stream: its work is ordered after
 *    work queued earlier on the legacy default (NULL) stream.  Callers that produce
 *    inputs on another stream pass that stream to aok_create or synchronise it first.  One context must not be used by two host threads at once;
 *    separate contexts are independent.
 */
#ifndef AOK_H
#define AOK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aok_context *aok_ctx;

typedef enum {
    AOK_OK = 0,
    AOK_ERR_ARG = 1,      /* invalid argument / shape mismatch (reference: ErrorException, AssertionError) */
    AOK_ERR_SPECIES = 2,  /* species index outside 1..NZ (reference: "species ... not in basis species list") */
    AOK_ERR_LMAX = 3,     /* angular momentum beyond what the kernels are built for */
    AOK_ERR_DEVICE = 4,   /* HIP runtime failure */
    AOK_ERR_NOMEM = 5,    /* workspace allocation failed */
    AOK_ERR_UNSUPPORTED = 6
} aok_status;

typedef enum { AOK_F32 = 0, AOK_F64 = 1 } aok_dtype;
typedef enum { AOK_I64 = 0, AOK_I32 = 1 } aok_itype;
typedef enum { AOK_DECAY_GAUSSIAN = 0, AOK_DECAY_SLATER = 1 } aok_decay;

/* Overlap execution modes (flags). */
#define AOK_FAST 0u        /* shell-pair-major, species-pair-bucketed kernels (default) */
#define AOK_PLAIN 1u       /* one work item per (b, s_a, s_b): the reference's decomposition,
                              runtime l loops, private E scratch, per-axis exps, float atomics */

/* Largest angular momentum supported by the kernels. */
#define AOK_AO_LMAX 4      /* AO evaluation: harmonics up to l = 4 (g)          */
#define AOK_OVL_LMAX 4     /* overlaps: shells up to l = 4 (g), pullback to l+2 */

/* ---- context -------------------------------------------------------------------- */
/* No reference counterpart (the reference's KernelAbstractions backend is implicit in
 * its array types, src/AtomicOrbitalKernels.jl:20-22).  stream: a hipStream_t to
 * borrow, or NULL to create a private one on `device`. */
aok_status aok_create(aok_ctx *ctx, int device, void *stream);
aok_status aok_destroy(aok_ctx ctx);
const char *aok_last_error(aok_ctx ctx);
/* Stream the context enqueues on (hipStream_t). */
void *aok_stream(aok_ctx ctx);
/* Library version string. */
const char *aok_version(void);

/* ---- H1: atomic orbitals ----------------------------------------------------------- */
/* Replaces evaluate(basis::AtomicOrbitals, Rs, sidx, ps, st)  (src/orbitals/atomicorbitals.jl:148-162)
 * and, with dP != NULL, evaluate_ed(basis, Rs, sidx, ps, st)  (atomicorbitals.jl:192-210):
 *   Rs        -> X [3*nX] in the basis's length unit; lengthscale = basis.lengthscale
 *                (atomicorbitals.jl:50-54,138-141)
 *   ps.Rnl    -> zeta, D [nRad*K*NZ]                 (gtostoradials.jl:14-20)
 *   st.Rnl    -> poly [nRad] (host; NULL = all 0, the Gaussian case, gtostoradials.jl:103-106)
 *   st.iR/iY  -> iR, iY [nOrb] (host, 1-based)        (atomicorbitals.jl:56-75,90-98)
 *   L         -> SolidHarmonics(L) of basis.Ylm (0..AOK_AO_LMAX)
 *   decay     -> Gaussian (r^2) or Slater (r) radial family (gtostoradials.jl:182-191)
 *   P [nX*nOrb] output; dP [3*nX*nOrb] output or NULL.
 * dP is the gradient w.r.t. the input-unit coordinates (includes lengthscale, :208). */
aok_status aok_ao_eval(aok_ctx ctx, aok_dtype dtype,
                       const void *X, const void *sidx, aok_itype itype, int64_t nX,
                       const void *zeta, const void *D, int64_t nRad, int64_t K, int64_t NZ,
                       const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb,
                       int L, double lengthscale, aok_decay decay,
                       void *P, void *dP);

/* Replaces the rrule(evaluate, basis, X, ps, st) pullback (atomicorbitals.jl:307-328):
 *   dX    [3*nX]  <- _aorb_pullback_x(dP, evaluate_ed(...)[2])   (atomicorbitals.jl:283-299)
 *   dzeta, dD [nRad*K*NZ] <- pullback_ps(dP, basis, Rs, sidx, ps, st).Rnl
 *                            (atomicorbitals.jl:254-267, gtostoradials.jl:232-267)
 * Any of dX / dzeta / dD may be NULL to skip it.  Parameter cotangents are reduced in
 * FP64 in a fixed order (deterministic) and rounded to dtype. */
aok_status aok_ao_pullback(aok_ctx ctx, aok_dtype dtype,
                           const void *X, const void *sidx, aok_itype itype, int64_t nX,
                           const void *zeta, const void *D, int64_t nRad, int64_t K, int64_t NZ,
                           const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb,
                           int L, double lengthscale, aok_decay decay,
                           const void *dPin, void *dX, void *dzeta, void *dD);

/* ---- H2: Cartesian normalisation ---------------------------------------------------- */
/* Replaces _compile_coef(zeta, D, ls) (src/integrals/orbital/cartesian_gto_basis.jl:110-120):
 * coef = D / sqrt(normsq) per (shell, species), computed in FP64 on the device and
 * written in dtype.  ls [ns] host. */
aok_status aok_compile_coef(aok_ctx ctx, aok_dtype dtype, const void *zeta, const void *D,
                            int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, void *coef);

/* Replaces _compile_coef_pb(dcoef, zeta, D, ls) (overlap_2c_grad.jl:251-286), FP64 inside. */
aok_status aok_compile_coef_pb(aok_ctx ctx, aok_dtype dtype, const void *dcoef,
                               const void *zeta, const void *D, int64_t ns, int64_t K, int64_t NZ,
                               const int64_t *ls, void *dD, void *dzeta_norm);

/* ---- H2: 2-centre overlap ------------------------------------------------------------ */
/* Replaces batch_overlap!(out, basis::CartesianGTOBasis, XA, XB)
 * (src/integrals/orbital/overlap_2c.jl:171-197) once XA/XB have been unpacked to
 * positions + species (overlap_2c.jl:139-158):
 *   posA, posB [3*B] in the basis's length unit, scaled by lengthscale in dtype;
 *   sidxA, sidxB [B] 1-based; coef, alpha [ns*K*NZ] (basis.coef, basis.zeta);
 *   ls [ns] host; out [N*N*B], N = sum (l+1)(l+2)/2.
 * Output semantics are the reference's, including the N1 = 2l_a+1 contraction stride
 * and nbf_a read-back for l >= 2 (overlap_2c.jl:256-281) and the coef == 0 skip.
 * flags: AOK_FAST (default) or AOK_PLAIN. */
aok_status aok_ovl2c(aok_ctx ctx, aok_dtype dtype, unsigned flags,
                     const void *posA, const void *posB, const void *sidxA, const void *sidxB,
                     aok_itype itype, int64_t B,
                     int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                     const void *coef, const void *alpha, double lengthscale, void *out);

/* Replaces batch_overlap(orb, XA, XB, ps, st) (overlap_2c_grad.jl:41-57): as aok_ovl2c,
 * but coef is recompiled from ps.Rnl = (zeta, D) in FP64 on the device first
 * (_compiled_from_ps, overlap_2c_grad.jl:22-30). */
aok_status aok_ovl2c_ps(aok_ctx ctx, aok_dtype dtype, unsigned flags,
                        const void *posA, const void *posB, const void *sidxA, const void *sidxB,
                        aok_itype itype, int64_t B,
                        int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                        const void *zeta, const void *D, double lengthscale, void *out);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * forward of aok_ovl2c (fast kernels, flags must be 0), same arguments and results, with
 * its bucketing tables and primitive-pair records in a fixed per-device array of
 * AOK_NOALLOC_SCRATCH_BYTES that lives until the process ends, instead of the context
 * workspace.  It calls no hipMalloc/hipFree.  A batch whose tables exceed the array
 * (about 8 bytes per pair plus 32 bytes per (species pair, shell pair, primitive pair):
 * some 4 million pairs of a 3-species, 6-shell, 9-primitive basis) fails with
 * AOK_ERR_ARG before any launch; split it.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other contexts, other streams, other host
 * threads) race and must be serialised by the caller.  Untimed (aok_set_timing does not
 * apply). */
#define AOK_NOALLOC_SCRATCH_BYTES (32u << 20)
aok_status aok_ovl2c_noalloc(aok_ctx ctx, aok_dtype dtype, unsigned flags,
                             const void *posA, const void *posB,
                             const void *sidxA, const void *sidxB, aok_itype itype, int64_t B,
                             int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                             const void *coef, const void *alpha, double lengthscale, void *out);

/* Replaces _overlap_pb_coef_alpha(dS, basis, XA, XB) (overlap_2c_grad.jl:218-242):
 * dcoef, dalpha [ns*K*NZ] (dtype) from the cotangent dS [N*N*B].  Every primitive
 * slot receives its cotangent, padded (coef == 0) slots included
 * (overlap_2c_grad.jl:104-110).  Reductions in FP64, fixed order. */
aok_status aok_ovl2c_pullback(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS,
                              const void *posA, const void *posB, const void *sidxA,
                              const void *sidxB, aok_itype itype, int64_t B,
                              int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                              const void *coef, const void *alpha, double lengthscale,
                              void *dcoef, void *dalpha);

/* Replaces pullback_ps(dS, orb, XA, XB, ps, st) (overlap_2c_grad.jl:296-304), the body of
 * rrule(batch_overlap, ...) (:309-318): kernel pullback + normalisation VJP,
 * dzeta = dalpha + dzeta_norm, dD; FP64 inside, written in dtype. */
aok_status aok_ovl2c_pullback_ps(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS,
                                 const void *posA, const void *posB, const void *sidxA,
                                 const void *sidxB, aok_itype itype, int64_t B,
                                 int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                                 const void *zeta, const void *D, double lengthscale,
                                 void *dzeta, void *dD);

/* No reference counterpart: the reference's rrule returns NoTangent for the positions
 * (overlap_2c_grad.jl:309-318).  Position cotangents of the 2-centre overlap, for callers
 * that differentiate through geometry (forces, position-dependent losses):
 *   dA[3b+c] = sum_ij dS[i,j,b] dS_ij/dA_bc  (per unit of the input length unit),
 *   dB = -dA (S depends on A - B only); dB may be NULL.
 * The derivative is of the forward as computed (same alias map, zero coefficients
 * skipped); per-pair sums in FP64, fixed order.  Fast path only (AOK_PLAIN is rejected). */
aok_status aok_ovl2c_pullback_x(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS,
                                const void *posA, const void *posB, const void *sidxA,
                                const void *sidxB, aok_itype itype, int64_t B,
                                int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                                const void *coef, const void *alpha, double lengthscale,
                                void *dA, void *dB);

/* ---- H2: 2-centre plans (forward + pullback sharing one preparation) ----------------- */
/* rrule(batch_overlap, orb, XA, XB, ps, st) (overlap_2c_grad.jl:309-318) runs the forward
 * and, later, pullback_ps on the same batch and parameters; the reference recompiles the
 * coefficients and re-walks the shell pairs in each.  A plan does the per-(sidx, params)
 * work once — the FP64 coefficient compilation (_compiled_from_ps, :22-30), the
 * species-pair bucketing of the batch and the primitive-pair records — and both calls
 * reuse it.
 *   params = AOK_PARAMS_COEF: (p1, p2) = (coef, alpha), pullback writes (dcoef, dalpha)
 *            as aok_ovl2c_pullback;
 *   params = AOK_PARAMS_PS:   (p1, p2) = (zeta, D), pullback writes (dzeta, dD) as
 *            aok_ovl2c_pullback_ps.
 * The plan references sidxA/sidxB/p1/p2 (device) until it is destroyed: they must stay
 * valid and unchanged.  Positions are per call.  A plan is used on its context's stream
 * and, like the context, by one host thread at a time.  Results are identical to the
 * corresponding one-shot entry points. */
#define AOK_PARAMS_COEF 0
#define AOK_PARAMS_PS 1
typedef struct aok_ovl2c_plan_s *aok_ovl2c_plan;

aok_status aok_ovl2c_plan_create(aok_ctx ctx, aok_dtype dtype, unsigned flags,
                                 const void *sidxA, const void *sidxB, aok_itype itype,
                                 int64_t B, int64_t ns, int64_t K, int64_t NZ,
                                 const int64_t *ls, int params, const void *p1,
                                 const void *p2, double lengthscale, aok_ovl2c_plan *plan);
/* Forward: out [N*N*B] as aok_ovl2c / aok_ovl2c_ps. */
aok_status aok_ovl2c_plan_forward(aok_ovl2c_plan plan, const void *posA, const void *posB,
                                  void *out);
/* Pullback of dS [N*N*B]: (d1, d2) [ns*K*NZ] as described above. */
aok_status aok_ovl2c_plan_pullback(aok_ovl2c_plan plan, const void *dS, const void *posA,
                                   const void *posB, void *d1, void *d2);
/* Position cotangents of dS through the plan: (dA, dB) [3*B] as aok_ovl2c_pullback_x. */
aok_status aok_ovl2c_plan_pullback_x(aok_ovl2c_plan plan, const void *dS, const void *posA,
                                     const void *posB, void *dA, void *dB);
/* No reference counterpart: re-read the species indices and parameters the plan was
 * created from (same pointers and sizes, contents changed in place, e.g. by an optimiser
 * step) and rebuild its bucketing and primitive-pair records.  Captured in a graph, it
 * makes every replay see the parameters as they are at replay time. */
aok_status aok_ovl2c_plan_update(aok_ovl2c_plan plan);
aok_status aok_ovl2c_plan_destroy(aok_ovl2c_plan plan);

/* ---- H2: 3-centre overlap ------------------------------------------------------------ */
/* Replaces batch_overlap_3c!(out, basis::CompiledBasis, posA, posB, posC)
 * (src/integrals/kernels_3c.jl:183-200, kernel batch_V_kernel! :12-162), generalised to
 * per-centre species with the orbital path's padded [ns*K*NZ] primitive layout; with
 * NZ = 1 and NULL sidx it is the reference's single-basis form (a CompiledBasis with
 * nprim/prim_offset maps onto it by zero-coefficient padding).  posA/B/C [3*B] in
 * Bohr after lengthscale; out [N*N*N*B] with the N1, N1*N2 contraction strides. */
aok_status aok_ovl3c(aok_ctx ctx, aok_dtype dtype, unsigned flags,
                     const void *posA, const void *posB, const void *posC,
                     const void *sidxA, const void *sidxB, const void *sidxC,
                     aok_itype itype, int64_t B,
                     int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                     const void *coef, const void *alpha, double lengthscale, void *out);

/* No reference counterpart (the reference has no 3-centre pullback).  Parameter
 * cotangents of aok_ovl3c: dcoef, dalpha [ns*K*NZ] (dtype) from dT [N*N*N*B], with the
 * rules of aok_ovl2c_pullback (every primitive slot receives its coefficient cotangent,
 * padded slots included; exponent cotangents by the raised-momentum identity; FP64 sums
 * in a fixed order).  Shells up to l = 2 (AOK_ERR_LMAX beyond); flags are ignored (one
 * implementation).  The (zeta, D) cotangents of an orbital basis follow with
 * aok_compile_coef_pb: dzeta = dalpha + dzeta_norm. */
aok_status aok_ovl3c_pullback(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dT,
                              const void *posA, const void *posB, const void *posC,
                              const void *sidxA, const void *sidxB, const void *sidxC,
                              aok_itype itype, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                              const int64_t *ls, const void *coef, const void *alpha,
                              double lengthscale, void *dcoef, void *dalpha);

/* ---- HIP graphs ---------------------------------------------------------------------- */
/* No reference counterpart (the reference relaunches every kernel per call).  Between
 * aok_graph_begin and aok_graph_end the context's calls are recorded, not run: nothing
 * synchronises, no timing events are taken, and a call that would need a larger
 * workspace (run it once eagerly first) or plan buffer fails and fails the capture.
 * The graph keeps every device pointer it was recorded with: the caller keeps those
 * buffers and plans alive, and updates inputs in place, until aok_graph_destroy.
 * aok_graph_launch replays the whole recorded sequence as one launch on the context
 * stream; in blocking mode it waits and reports device-side errors like any call.
 * Workspaces the library outgrows while graphs are alive are kept until the last graph
 * of the context is destroyed.  Destroying the context first detaches its graphs: they
 * then only accept aok_graph_destroy. */
typedef struct aok_graph_s *aok_graph;
aok_status aok_graph_begin(aok_ctx ctx);
aok_status aok_graph_end(aok_ctx ctx, aok_graph *graph);
aok_status aok_graph_launch(aok_graph graph);
aok_status aok_graph_destroy(aok_graph graph);

/* ---- tuning -------------------------------------------------------------------------------- */
/* No reference counterpart (the reference has no tunables): set / read a per-context
 * launch-shape knob used by the kernels and by the autotuner.  Keys:
 *   "ao_ppl"   AO values-only evaluation: consecutive points per lane (1, 2 or 4; 0 = default)
 *   "ao_grid"  AO kernels: persistent grid size in workgroups (0 = default)
 *   "ovl_cap"  2-centre kernels: LDS stripe capacity in elements per lane (0 = default: 40
 *              for the forward, 28 for the pullback and position cotangents)
 *   "ovl_cap_pb" the same for the pullback and position cotangents only (0 = ovl_cap)
 *   "ovl_split" 2-centre pullback / position cotangents of bases with f or g shells: run
 *              the chunks of s, p, d shells in the kernel instance compiled for l <= 2
 *              (0 = default = 1; 2 = every chunk in the instance for the largest l)
 *   "ovl_order" 2-centre fast kernels' workgroup order (every tile's workgroups run on one
 *              XCD): 1 tile-major (consecutive workgroups take consecutive tiles of one
 *              stripe chunk), 2 XCD-grouped (a tile's chunks one after another, sharing the
 *              XCD's L2), k >= 3 groups of k - 1 chunks; 0 = default: 2 when N > 32, else 1
 *   "ovl3_parts" 3-centre pullback: split each triple's shell loops into parts run by
 *              separate threads (each with its own FP64 partial row): 1 none, 2 by
 *              third-centre shell, 3 by (third, second) shell pair; 0 = default: the finest
 *              split whose rows for min(B, 4096) triples fit 1 GiB; split rows are processed
 *              in batch chunks that fit 1 GiB
 * Unknown keys or out-of-range values return AOK_ERR_ARG. */
aok_status aok_tune(aok_ctx ctx, const char *key, int64_t value);
aok_status aok_tune_get(aok_ctx ctx, const char *key, int64_t *value);

/* ---- measurement hooks ----------------------------------------------------------------- */
/* Device time (ms) of the dominant kernel of the most recent call on this context, taken
 * with HIP events on the context stream around that kernel's launch (0 if none; in
 * stream-ordered mode the most recent timed launch, waited for). */
double aok_last_kernel_ms(aok_ctx ctx);
/* Name of that kernel. */
const char *aok_last_kernel_name(aok_ctx ctx);
/* Every timed launch not yet returned, oldest first: waits for them, writes up to `cap`
 * (name, ms) pairs and removes those from the record; *count = pairs written.  Names
 * are static strings.  The record keeps the latest 65536 launches. */
aok_status aok_kernel_times(aok_ctx ctx, int64_t cap, double *ms, const char **names, int64_t *count);
/* Enable/disable the per-call kernel timing events (default off).  A timed call waits
 * for its work in either execution mode. */
aok_status aok_set_timing(aok_ctx ctx, int enable);

/* ---- execution mode ---------------------------------------------------------------------- */
/* No reference counterpart.  Blocking (the default, and the reference's semantics: every
 * wrapper ends in KA.synchronize) -- each entry point returns after its work has finished
 * and reports device-detected errors (species index out of range) itself.
 * Stream-ordered (blocking = 0) -- each entry point returns once its work is enqueued on
 * the context stream; argument errors are still reported at once, device-detected errors
 * are latched and reported by aok_sync (or when blocking mode is switched back on).
 * Outputs may be read only after the stream has reached them; inputs must stay valid and
 * unchanged until then.  Leaving stream-ordered mode synchronises. */
aok_status aok_set_blocking(aok_ctx ctx, int blocking);
/* 1 blocking, 0 stream-ordered, -1 for a NULL context. */
int aok_get_blocking(aok_ctx ctx);
/* Wait for all work enqueued by this context; report and clear latched device errors. */
aok_status aok_sync(aok_ctx ctx);

#ifdef __cplusplus
}
#endif
#endif /* AOK_H */
