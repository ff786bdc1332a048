/*
 * qk_abi.h - ABI constants shared by every entry point of the MI355X
 * Monte Carlo pricing library.
 *
 * The values and the QKPluginAPI layout are the reference's
 * (cpp/include/quantkernel/qk_abi.h:11-31) so that an existing caller that
 * negotiates ABI 2.x (x >= 12) keeps working.  The minor version is 13: the
 * additions of this library (standard-error outputs, RNG mode selection,
 * mixed-model batches, LSMC basis selection, device-resident batches and the
 * Heston surrogate) are new names only; no 2.12 signature changes.
 */
#ifndef QK_ABI_H
#define QK_ABI_H

#ifdef __cplusplus
#include <cstdint>
#else
#include <stdint.h>
#endif

#define QK_ABI_MAJOR 2
#define QK_ABI_MINOR 13

/* Return codes (reference qk_abi.h:15-20). */
#define QK_OK                 0
#define QK_ERR_NULL_PTR      -1
#define QK_ERR_BAD_SIZE      -2
#define QK_ERR_ABI_MISMATCH  -3
#define QK_ERR_RUNTIME_INIT  -4   /* no usable gfx950 device, or a HIP call failed */
#define QK_ERR_INVALID_INPUT -5

/* Option type (reference qk_abi.h:22-24). */
#define QK_CALL 0
#define QK_PUT  1

/* Normal-stream selector for the Monte Carlo entry points (ABI 2.13). */
#define QK_RNG_PHILOX        0  /* on-device Philox4x32-10 + Box-Muller, FP32 paths (default) */
#define QK_RNG_MT19937_FP64  1  /* host mt19937_64 + std::normal_distribution stream, FP64 paths */
#define QK_RNG_MT19937_FP32  2  /* same host stream, FP32 paths (reports FP32 path error) */

/* Model ids of the mixed-model batch (ABI 2.13). */
#define QK_MODEL_GBM_EULER     0
#define QK_MODEL_GBM_MILSTEIN  1
#define QK_MODEL_HESTON_EULER  2
#define QK_MODEL_LOCAL_VOL     3   /* constant local vol, Euler */

/* Longstaff-Schwartz regression basis (ABI 2.13). */
#define QK_LSMC_BASIS_REFERENCE 0  /* {1, S, S^2}, 3x3 Gauss-Jordan, |pivot| < 1e-14 -> beta = 0 */
#define QK_LSMC_BASIS_CUBIC     1  /* {1, x, x^2, x^3} with x = S/K, 4x4 FP64 normal equations */

typedef struct QKPluginAPI {
    int32_t     abi_major;
    int32_t     abi_minor;
    const char* plugin_name;
} QKPluginAPI;

#endif /* QK_ABI_H */
