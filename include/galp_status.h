/* galp_status.h -- status codes shared by the C-ABI (galp.h, galp_host.h).
 * The reference ignores every CUDA status and only calls exit(1)
 * (label_propagation/src/main.cu:47,62); every galp_* entry point returns one
 * of these instead. */
#ifndef GALP_STATUS_H
#define GALP_STATUS_H

#ifdef __cplusplus
extern "C" {
#endif

/* 0 = success; negative values are errors.  No entry point throws. */
typedef enum galp_status {
    GALP_OK = 0,
    GALP_EINVAL = -1,      /* bad argument or malformed graph */
    GALP_ENOMEM = -2,      /* device or host memory, or the caller's budget */
    GALP_EHIP = -3,        /* a HIP runtime call failed */
    GALP_EUNSUPPORTED = -4,/* policy/method outside the one-GPU build */
    GALP_EIO = -5,         /* file missing, truncated or malformed */
    GALP_ESTATE = -6       /* call out of order (e.g. step before init) */
} galp_status;

#ifdef __cplusplus
}
#endif
#endif
