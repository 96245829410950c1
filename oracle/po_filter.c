/* Oracle: the supertensor interleave filter.  TEST INFRASTRUCTURE ONLY (see po.h).
 *
 * pentago/data/filter.h (interleave_super / uninterleave_super): for word k of the
 * (black, white) supers, output word 2k holds bits 0..31 and word 2k+1 bits 32..63 of both,
 * black's bit i at bit 2i and white's at bit 2i+1.  Restated bit by bit from that
 * definition. */
#include "po.h"

void po_interleave(int64_t n, po_super* nodes) {
  for (int64_t j = 0; j < n; j++) {
    const po_super b = nodes[2 * j], w = nodes[2 * j + 1];
    uint64_t out[8] = {0};
    for (int k = 0; k < 4; k++)
      for (int i = 0; i < 64; i++) {
        const int word = 2 * k + (i >> 5), bit = 2 * (i & 31);
        out[word] |= (b.w[k] >> i & 1) << bit | (w.w[k] >> i & 1) << (bit + 1);
      }
    for (int k = 0; k < 4; k++) { nodes[2 * j].w[k] = out[k]; nodes[2 * j + 1].w[k] = out[4 + k]; }
  }
}

void po_uninterleave(int64_t n, po_super* nodes) {
  for (int64_t j = 0; j < n; j++) {
    uint64_t in[8];
    for (int k = 0; k < 4; k++) { in[k] = nodes[2 * j].w[k]; in[4 + k] = nodes[2 * j + 1].w[k]; }
    po_super b = {{0}}, w = {{0}};
    for (int k = 0; k < 4; k++)
      for (int i = 0; i < 64; i++) {
        const int word = 2 * k + (i >> 5), bit = 2 * (i & 31);
        b.w[k] |= (in[word] >> bit & 1) << i;
        w.w[k] |= (in[word] >> (bit + 1) & 1) << i;
      }
    nodes[2 * j] = b;
    nodes[2 * j + 1] = w;
  }
}
