/* Error codes returned by every C entry point of the library.
 * The numeric values are part of the interface: they are the ones of the bipp 1.0.0
 * BippError enumeration (include/bipp/errors.h:5-66 of the reference), so code that
 * switches on them keeps working. */
#pragma once

enum BippError {
  BIPP_SUCCESS = 0,
  BIPP_GENERIC_ERROR = 1,
  BIPP_INTERNAL_ERROR = 2,
  BIPP_INVALID_PARAMETER_ERROR = 3,
  BIPP_INVALID_POINTER_ERROR = 4,
  BIPP_INVALID_HANDLE_ERROR = 5,
  BIPP_EIGENSOLVER_ERROR = 6,
  BIPP_NOT_IMPLEMENTED_ERROR = 7,
  BIPP_GPU_ERROR = 8,
  BIPP_GPU_SUPPORT_ERROR = 9,
  BIPP_GPU_BLAS_ERROR = 10,
  BIPP_INVALID_ALLOCATOR_FUNCTION = 11,
  BIPP_MPI_ERROR = 12,
  BIPP_HDF5_ERROR = 13,
  BIPP_FILE_ERROR = 14
};

#ifndef __cplusplus
typedef enum BippError BippError;
#endif
