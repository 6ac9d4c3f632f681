/*
 * Status codes written to the `ierr` argument of every hmat_lib call.
 * Same enumerators, same order (so the same integer values) as the
 * reference (include/hmat_lib/error.h:9-16).  Device-side failures of the
 * HIP layer are reported through these codes as well: out-of-memory and
 * missing-device errors as ALLOCATION_FAILURE, kernel-side SVD problems as
 * SVD_FAILURE, QR problems as QR_FAILURE.  No new codes are added.
 */
#ifndef HODLR_ERROR_H
#define HODLR_ERROR_H

enum ErrorCode {
  SUCCESS,
  ALLOCATION_FAILURE,
  SVD_FAILURE,
  SVD_ALLOCATION_FAILURE,
  INPUT_ERROR,
  QR_FAILURE
};

#endif
