// Launch geometry for gfx950.  The reference keeps an sm_86 table of
// threads x items per thread by key/value width
// (include/rsort/detail/tuning.cuh:138-213); this is the wave64 / 160 KiB-LDS
// counterpart.  A tile of THREADS*ITEMS keys is one workgroup of the pass
// kernel; the staging buffer holds one tile of keys (then of values), so the
// LDS footprint is THREADS*ITEMS*max(key,value bytes) + 1 KiB per wave.
#pragma once

#include <stdint.h>

#include "rsort_protocol.hpp"

namespace rsa {

constexpr int kHistThreads = 512;

// LDS copies of the per-pass digit counters in the histogram kernel: one per
// wave while it fits in 64 KiB, otherwise shared by pairs of waves.
RSA_HD constexpr int hist_copies(int key_bytes) {
    return (key_bytes * 1024 * (kHistThreads / 64) <= 64 * 1024) ? (kHistThreads / 64)
                                                                   : (64 * 1024) / (key_bytes * 1024);
}

struct TileCfg {
    uint32_t threads;
    uint32_t items;
};

// key_bytes in {1,2,4,8,16}; val_bytes in {0,1,2,4,8,16}
RSA_HD constexpr TileCfg tile_cfg(int key_bytes, int val_bytes) {
    const int wide = key_bytes > val_bytes ? key_bytes : val_bytes;
    if (key_bytes <= 4) {
        if (val_bytes == 0) return TileCfg{512, 16};
        if (wide <= 4) return TileCfg{512, 16};
        if (wide == 8) return TileCfg{512, 12};
        return TileCfg{512, 8};
    }
    if (key_bytes == 8) {
        if (val_bytes == 0) return TileCfg{512, 12};
        if (wide == 8) return TileCfg{512, 10};
        return TileCfg{512, 8};
    }
    // 16-byte keys
    if (val_bytes == 0) return TileCfg{512, 8};
    return TileCfg{512, 6};
}

}  // namespace rsa
