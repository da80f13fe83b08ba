// sw2_batch.h -- the launchers of the batch kernels of the mixed-precision
// step, one per tile shape (sw2_batch.hip, compiled once per shape), called
// by the batch code of sw2_device.hip.  Internal to libsw2.
//
// Each queues one step's height and momentum kernels of a batch on stream s:
// mem and st are the device member records and step states, hdr the host copy
// of the header (the kernels take it by value); sch is
// SW2's limiter (0-9); hv is 0 (no hyperviscosity) or 2 (del-2 with del-4);
// tr selects the kernels that carry the tracer's code; ge and gu are the
// workgroups of the two kernels; pe and pu ask for persistent workgroups
// (32 x 16 tiles only).
#ifndef SW2_BATCH_H
#define SW2_BATCH_H

#include <hip/hip_runtime.h>

#define SW2_BATCH_LAUNCHER(name)                                                                                      \
    __attribute__((visibility("hidden"))) void name(const void *mem, const void *st, const void *hdr, int sch, int hv, \
                                                    int tr, int ge, int gu, int pe, int pu, hipStream_t s)

SW2_BATCH_LAUNCHER(sw2_batch_phases_32x16);
SW2_BATCH_LAUNCHER(sw2_batch_phases_64x8);
SW2_BATCH_LAUNCHER(sw2_batch_phases_64x16);

#endif
