/*
 * beagle_amd_ext.h -- extension entry points of libhmsbeagle_amd.so that
 * have no counterpart in the reference API (libhmsbeagle/beagle.h).  They
 * add to beagle.h without changing any reference signature.
 */
#ifndef BEAGLE_AMD_EXT_H
#define BEAGLE_AMD_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Returns the HIP stream (hipStream_t) on which the instance enqueues all
 * of its work, so a caller can order its own work after it or bracket it
 * with events.  No counterpart in the reference. */
int beagleExtGetStream(int instance, void** outStream);

#ifdef __cplusplus
}
#endif

#endif /* BEAGLE_AMD_EXT_H */
