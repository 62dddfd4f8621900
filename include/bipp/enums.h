/* Enumerations shared by the C and C++ interfaces.  Values match bipp 1.0.0
 * (include/bipp/enums.h:5-15 of the reference). */
#pragma once

enum BippProcessingUnit { BIPP_PU_AUTO = 0, BIPP_PU_CPU = 1, BIPP_PU_GPU = 2 };

enum BippPrecision { BIPP_PRECISION_SINGLE = 0, BIPP_PRECISION_DOUBLE = 1 };

enum BippLogLevel {
  BIPP_LOG_LEVEL_OFF = 0,
  BIPP_LOG_LEVEL_ERROR = 1,
  BIPP_LOG_LEVEL_WARN = 2,
  BIPP_LOG_LEVEL_INFO = 3,
  BIPP_LOG_LEVEL_DEBUG = 4
};

/* Arithmetic used by the fused GPU pipeline (no counterpart in the reference, which
 * has a single precision switch):
 *   BIPP_COMPUTE_FP64      -- everything in float64 (the GPU oracle form);
 *   BIPP_COMPUTE_FP32      -- everything in float32, phases in float32 (plain form);
 *   BIPP_COMPUTE_MIXED     -- float64 eigensolve and phase range reduction, image
 *                             contraction on split-bf16 (hi+lo) MFMA, float32 accumulate. */
enum BippCompute { BIPP_COMPUTE_FP64 = 0, BIPP_COMPUTE_FP32 = 1, BIPP_COMPUTE_MIXED = 2 };

/* Eigenvalue filters of python/bipp/filter.py:7-26. */
enum BippFilter {
  BIPP_FILTER_LSQ = 0,
  BIPP_FILTER_STD = 1,
  BIPP_FILTER_SQRT = 2,
  BIPP_FILTER_INV = 3,
  BIPP_FILTER_INV_SQRT = 4
};

#ifndef __cplusplus
typedef enum BippProcessingUnit BippProcessingUnit;
typedef enum BippLogLevel BippLogLevel;
typedef enum BippPrecision BippPrecision;
typedef enum BippCompute BippCompute;
typedef enum BippFilter BippFilter;
#endif
