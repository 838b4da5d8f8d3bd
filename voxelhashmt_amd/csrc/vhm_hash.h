// Find-or-insert of block keys in the persistent open-addressing table
// (device side; shared by the frame allocator and the point-cloud path).
#pragma once

#include "vhm_internal.h"

namespace vhm {

// Inserts key into the persistent table; returns its slot (kInvalid if the
// probe bound is exhausted).  `created` is set for the inserting lane.
__device__ inline uint32_t findOrInsert(const VolumeView& V, uint64_t key, bool& created) {
  uint32_t slot = hashSlot(key, V.log2cap);
  const int bound = (V.mask + 1u) < static_cast<uint32_t>(kMaxProbe) ? static_cast<int>(V.mask + 1u) : kMaxProbe;
  for (int probe = 0; probe < bound; ++probe) {
    // A plain (cacheable) load: a slot's key only ever changes from empty to
    // its key, so a stale read can only show "empty", which the CAS below
    // then resolves.
    const uint64_t cur = V.keys[slot];
    if (cur == key) return slot;
    if (cur == kEmptyKey) {
      const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&V.keys[slot]),
                                      static_cast<unsigned long long>(kEmptyKey),
                                      static_cast<unsigned long long>(key));
      if (prev == kEmptyKey) {
        created = true;
        return slot;
      }
      if (prev == key) return slot;
    }
    slot = (slot + 1u) & V.mask;
  }
  return kInvalid;
}

// The block of key: its table slot, taking a zeroed pool slot when the key
// is new (blocks_[key], voxel_hash_tsdf.h:133-145).  kInvalid (with the
// error bit set) when the table's probe bound is exhausted; a full pool sets
// kErrPoolFull and leaves the slot's pool index unset.
__device__ inline uint32_t insertBlock(const VolumeView& V, uint64_t key) {
  bool created = false;
  const uint32_t slot = findOrInsert(V, key, created);
  if (slot == kInvalid) {
    atomicOr(&V.hdr->error, kErrHashFull);
    return kInvalid;
  }
  if (created) {
    const uint32_t p = atomicAdd(&V.hdr->num_blocks, 1u);
    if (p < V.max_blocks) {
      V.vals[slot] = p;
      V.pool_keys[p] = key;
    } else {
      atomicOr(&V.hdr->error, kErrPoolFull);
    }
  }
  return slot;
}

}  // namespace vhm
