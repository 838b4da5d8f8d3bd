// The persistent open-addressing table of block keys (device side): lookup,
// find-or-insert, and the ordered numbering of each pass's new blocks.
#pragma once

#include "vhm_internal.h"

namespace vhm {

// Table slot of key, kInvalid when absent (linear probing, as findOrInsert).
__device__ inline uint32_t findSlot(const VolumeView& V, uint64_t key) {
  uint32_t slot = hashSlot(key, V.log2cap);
  const int bound = (V.mask + 1u) < static_cast<uint32_t>(kMaxProbe) ? static_cast<int>(V.mask + 1u) : kMaxProbe;
  for (int probe = 0; probe < bound; ++probe) {
    const uint64_t cur = V.keys[slot];
    if (cur == key) return slot;
    if (cur == kEmptyKey) return kInvalid;
    slot = (slot + 1u) & V.mask;
  }
  return kInvalid;
}

// Pool index of block (x, y, z), kInvalid when it is not allocated.
__device__ inline uint32_t lookupBlock(const VolumeView& V, int x, int y, int z) {
  if (!keyInRange(x, y, z)) return kInvalid;
  const uint32_t slot = findSlot(V, packKey(x, y, z));
  if (slot == kInvalid) return kInvalid;
  const uint32_t p = V.vals[slot];
  return p < V.max_blocks ? p : kInvalid;
}

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
// kErrPoolFull and leaves the slot's pool index unset.  The pool index taken
// here depends on the order in which lanes arrive; launchNumberBlocks
// renumbers the pass's new blocks in key order before anything reads them.
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
