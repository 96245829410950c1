// Host-side lookup tables for the MI355X engine.
//
// The tables are the ones the reference bakes in at build time
// (pentago/base/precompute.cc); this file regenerates them at library init from their
// mathematical definitions and checks each one against the SHA-1 digest the reference
// records for it, so a wrong table fails loudly before any kernel runs.
#pragma once
#include <cstdint>
#include <string>

namespace pe {

struct Tables {
  uint16_t pack[512];                 // side quadrant (9 bits) -> radix-3 digits of 1
  uint16_t unpack[19683][2];          // radix-3 quadrant -> (side0, side1)
  uint16_t rot[512][2];               // rotate a side quadrant left / right by 90 degrees
  uint16_t refl[512];                 // reflect a side quadrant about x = y
  uint64_t superwin[4 * 512 * 5 * 4]; // (quadrant, state, field) -> 256-bit super
  uint8_t commute[8][256];            // local id of g^-1 l g
  uint16_t rmin_offsets[56];          // rotation-minimal quadrants, bucketed by (black, white)
  uint16_t rmin_flat[19683];
  int rmin_flat_size;
  uint16_t rmin_inverse[19683];       // q -> 4 * index + rotation
  uint16_t rmin_moved[55];            // bucket entries moved by reflection
  uint64_t win_contrib[4][512];       // 2-bit counters of win patterns per quadrant state
};

// Build once (thread safe); throws std::runtime_error if a table digest mismatches.
const Tables& tables();

// Bucket of rotation-minimal quadrants with the given counts (pentago/base/section.cc:20-28).
int rmin_bucket(int black, int white);
int rmin_count(int black, int white);
const uint16_t* rmin_list(int black, int white);

std::string sha1_hex(const void* data, size_t n);
// Digest of a named table, hashed the way the reference hashes it.
std::string table_digest(const std::string& name);

}  // namespace pe
