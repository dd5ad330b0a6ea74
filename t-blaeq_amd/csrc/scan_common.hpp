// Decoupled look-back shared by the single-pass scans.
#pragma once
#include "device_common.hpp"

namespace tblaeq {

// Exclusive prefix of (count, weight) over all tiles before `tile`, read from
// the predecessors' status words by one wave (64 predecessors per probe).
__device__ void lookback2(const uint64_t* status, uint32_t epoch, uint64_t tile, int lane, uint64_t& ex0,
                          uint64_t& ex1) {
    uint64_t acc0 = 0, acc1 = 0;
    int64_t base = static_cast<int64_t>(tile) - 1;
    for (;;) {
        const int64_t t = base - lane;
        int s = 2;
        uint64_t v0 = 0, v1 = 0;
        if (t >= 0) {
            const uint64_t* st = status + static_cast<uint64_t>(t) * 4;
            const uint64_t i0 = ld_acquire(st + 2), i1 = ld_acquire(st + 3);
            if (status_valid(i0, epoch) && status_valid(i1, epoch)) {
                v0 = i0 & kValMask;
                v1 = i1 & kValMask;
            } else {
                const uint64_t a0 = ld_acquire(st + 0), a1 = ld_acquire(st + 1);
                if (status_valid(a0, epoch) && status_valid(a1, epoch)) {
                    s = 1;
                    v0 = a0 & kValMask;
                    v1 = a1 & kValMask;
                } else {
                    s = 0;
                }
            }
        }
        const uint64_t m2 = __ballot(s == 2);
        const uint64_t m0 = __ballot(s == 0);
        const int f2 = m2 ? __ffsll(static_cast<unsigned long long>(m2)) - 1 : 64;
        const int f0 = m0 ? __ffsll(static_cast<unsigned long long>(m0)) - 1 : 64;
        if (f0 < f2) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        acc0 += wave_sum_u64(lane <= f2 ? v0 : 0);
        acc1 += wave_sum_u64(lane <= f2 ? v1 : 0);
        if (f2 < 64) break;
        base -= 64;
    }
    ex0 = acc0;
    ex1 = acc1;
}


// Publishes a tile's (count, weight) aggregate and returns the exclusive prefix
// of all earlier tiles; called by wave 0 of the workgroup.  status holds four
// words per tile: aggregate pair, inclusive pair.
__device__ __forceinline__ void tile_prefix2(uint64_t* status, uint32_t epoch, uint64_t tile, int lane, uint64_t c,
                                             uint64_t s, uint64_t& ex0, uint64_t& ex1) {
    uint64_t* st = status + tile * 4;
    ex0 = 0;
    ex1 = 0;
    if (tile == 0) {
        if (lane == 0) {
            st_release(st + 2, status_word(epoch, c));
            st_release(st + 3, status_word(epoch, s));
        }
        return;
    }
    if (lane == 0) {
        st_release(st + 0, status_word(epoch, c));
        st_release(st + 1, status_word(epoch, s));
    }
    lookback2(status, epoch, tile, lane, ex0, ex1);
    if (lane == 0) {
        st_release(st + 2, status_word(epoch, ex0 + c));
        st_release(st + 3, status_word(epoch, ex1 + s));
    }
}

// Epoch of the current scan launch.  Every workgroup reads it when it starts;
// the last workgroup to leave advances it, so launches need no host-provided
// epoch and can be captured in a graph and replayed.
__device__ __forceinline__ uint32_t launch_epoch(ScanCtrl* ctrl) {
    return static_cast<uint32_t>(ld_acquire(&ctrl->seq) & 0x7FFFFFF);
}

// Called by thread 0 of every workgroup after its last tile: the last one to
// leave resets the ticket counters and advances the epoch.
__device__ __forceinline__ void scan_exit(ScanCtrl* ctrl) {
    const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(&ctrl->exitCounter), 1ull);
    if (old == gridDim.x - 1) {
        const uint64_t next = (ld_acquire(&ctrl->seq) + 1) & 0x7FFFFFF;
        st_release(&ctrl->seq, next);
        st_release(&ctrl->tileCounter, 0);
        st_release(&ctrl->exitCounter, 0);
    }
}

}  // namespace tblaeq
