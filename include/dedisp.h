/*
 * dedisp.h -- dedisp-compatible plan API of the MI355X (gfx950) dedisperser.
 *
 * Drop-in for the external dedisp library that casm-hella locates with
 * cmake/FindDEDISP.cmake:14-28 and calls at
 *   src/apps/casm_hella_refactored.cu:205-210   (create plan, DM list, queries)
 *   src/hella/src/dedisperse.cpp:32-38          (dedisp_execute_adv)
 *   src/hella/src/dedisperse.cpp:50-91          (the 17 dedisp_error names)
 * Names, types and argument meanings follow the public dedisp interface used
 * at those call sites.  Implemented by libhella_amd.so; every entry point
 * returns an error code and never exits the process.
 *
 * Semantics (all exact integer arithmetic on 8-bit input):
 *   out[d][t] = sum over live channels c of in[t + delay(d,c)][c]
 *   for t < nsamps - max_delay, with delay(d,c) = round_half_up(dm[d]*table[c])
 *   and table[c] = 4.15e3/dt * (1/(f0+c*df)^2 - 1/f0^2).
 *   out_nbits == 32 writes float(sum) (exact: sums stay below 2^24).
 */
#ifndef HELLA_AMD_DEDISP_H
#define HELLA_AMD_DEDISP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned int dedisp_word;
typedef unsigned long dedisp_size;
typedef int dedisp_bool;
typedef float dedisp_float;
typedef unsigned char dedisp_byte;

typedef struct dedisp_plan_struct *dedisp_plan;

/* src/hella/src/dedisperse.cpp:54-87 lists these names. */
typedef enum {
  DEDISP_NO_ERROR = 0,
  DEDISP_MEM_ALLOC_FAILED,
  DEDISP_MEM_COPY_FAILED,
  DEDISP_NCHANS_EXCEEDS_LIMIT,
  DEDISP_INVALID_PLAN,
  DEDISP_INVALID_POINTER,
  DEDISP_INVALID_STRIDE,
  DEDISP_NO_DM_LIST_SET,
  DEDISP_TOO_FEW_NSAMPS,
  DEDISP_INVALID_FLAG_COMBINATION,
  DEDISP_UNSUPPORTED_IN_NBITS,
  DEDISP_UNSUPPORTED_OUT_NBITS,
  DEDISP_INVALID_DEVICE_INDEX,
  DEDISP_DEVICE_ALREADY_SET,
  DEDISP_PRIOR_GPU_ERROR,
  DEDISP_INTERNAL_GPU_ERROR,
  DEDISP_UNKNOWN_ERROR
} dedisp_error;

/* flags; hella passes DEDISP_DEVICE_POINTERS (1<<2) at dedisperse.cpp:31 */
typedef enum {
  DEDISP_USE_DEFAULT = 0,
  DEDISP_HOST_POINTERS = 1 << 1,
  DEDISP_DEVICE_POINTERS = 1 << 2,
  DEDISP_WAIT = 1 << 3,
  DEDISP_ASYNC = 1 << 4
} dedisp_flag;

/* casm:205 -- dt in seconds, f0 and df in MHz (df < 0: descending band) */
dedisp_error dedisp_create_plan(dedisp_plan *plan, dedisp_size nchans,
                                dedisp_float dt, dedisp_float f0, dedisp_float df);
void dedisp_destroy_plan(dedisp_plan plan);

/* casm:207 -- ti is the intrinsic pulse width in microseconds */
dedisp_error dedisp_generate_dm_list(dedisp_plan plan, dedisp_float dm_start,
                                     dedisp_float dm_end, dedisp_float ti,
                                     dedisp_float tol);
dedisp_error dedisp_set_dm_list(dedisp_plan plan, const dedisp_float *dm_list,
                                dedisp_size count);
/* killmask[c] != 0 keeps channel c; NULL restores all channels */
dedisp_error dedisp_set_killmask(dedisp_plan plan, const dedisp_bool *killmask);
dedisp_error dedisp_set_gulp_size(dedisp_plan plan, dedisp_size gulp_size);
dedisp_size dedisp_get_gulp_size(dedisp_plan plan);

/* casm:208-210 */
const dedisp_float *dedisp_get_dm_list(const dedisp_plan plan);
dedisp_size dedisp_get_dm_count(const dedisp_plan plan);
dedisp_size dedisp_get_max_delay(const dedisp_plan plan);
dedisp_size dedisp_get_dm_delay(const dedisp_plan plan, int dm_trial);
dedisp_size dedisp_get_channel_count(const dedisp_plan plan);
const dedisp_bool *dedisp_get_killmask(const dedisp_plan plan);
dedisp_float dedisp_get_dt(const dedisp_plan plan);
dedisp_float dedisp_get_f0(const dedisp_plan plan);
dedisp_float dedisp_get_df(const dedisp_plan plan);

/* src/hella/src/dedisperse.cpp:32-38.  in: nsamps rows of in_stride bytes
 * (time-major, one byte per channel); out: dm_count rows of out_stride bytes,
 * nsamps - max_delay samples each.  Only in_nbits == 8 and out_nbits == 32
 * are supported. */
dedisp_error dedisp_execute_adv(const dedisp_plan plan, dedisp_size nsamps,
                                const dedisp_byte *in, dedisp_size in_nbits,
                                dedisp_size in_stride, dedisp_byte *out,
                                dedisp_size out_nbits, dedisp_size out_stride,
                                unsigned flags);
dedisp_error dedisp_execute(const dedisp_plan plan, dedisp_size nsamps,
                            const dedisp_byte *in, dedisp_size in_nbits,
                            dedisp_byte *out, dedisp_size out_nbits,
                            unsigned flags);
/* DM sub-range [first_dm_idx, first_dm_idx + dm_count) of the plan's list */
dedisp_error dedisp_execute_guru(const dedisp_plan plan, dedisp_size nsamps,
                                 const dedisp_byte *in, dedisp_size in_nbits,
                                 dedisp_size in_stride, dedisp_byte *out,
                                 dedisp_size out_nbits, dedisp_size out_stride,
                                 dedisp_size first_dm_idx, dedisp_size dm_count,
                                 unsigned flags);

/* No counterpart in the reference or in dedisp: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  dedisp_execute_adv with device pointers only
 * (DEDISP_DEVICE_POINTERS), allocating nothing: its scratch (descriptor, tile
 * table, the plan's delay tables and the channel-major copy of the input) is
 * one fixed 64 MiB device array per GPU, held until the process ends.  The
 * plan's host tables are built by the DM-list and killmask setters.
 * Requests needing more scratch than 64 MiB (about nchans * (nsamps + 2048)
 * bytes plus the delay tables) return DEDISP_MEM_ALLOC_FAILED with the reason
 * in hella_last_error().  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads, DEDISP_ASYNC calls
 * still running) race and must be serialised by the caller. */
dedisp_error dedisp_execute_adv_noalloc(const dedisp_plan plan, dedisp_size nsamps,
                                        const dedisp_byte *in, dedisp_size in_nbits,
                                        dedisp_size in_stride, dedisp_byte *out,
                                        dedisp_size out_nbits, dedisp_size out_stride,
                                        unsigned flags);

dedisp_error dedisp_set_device(int device_idx);
/* waits for every dedisp launch on the calling thread's plans */
dedisp_error dedisp_sync(void);
const char *dedisp_get_error_string(dedisp_error error);

/* Extension (no dedisp counterpart): run the plan's kernels on a caller's
 * hipStream_t (passed as void*); NULL selects the default stream. */
dedisp_error dedisp_set_stream(dedisp_plan plan, void *stream);

#ifdef __cplusplus
}
#endif
#endif
