/* Status codes shared by every entry point of the MI355X ALS engine. */
#ifndef ALS_STATUS_H
#define ALS_STATUS_H

#define ALS_OK 0
#define ALS_E_INVALID -1   /* bad argument, shape or configuration */
#define ALS_E_HIP -2       /* HIP runtime failure (allocation, launch, copy) */
#define ALS_E_IO -3        /* file missing, empty, wrong size or unreadable */
#define ALS_E_STATE -4     /* call out of order (e.g. update before set_train) */
#define ALS_E_NUMERIC -5   /* a factorisation broke down (non-SPD system) */
#define ALS_E_UNSUPPORTED -6 /* valid request this build does not provide */

#endif
