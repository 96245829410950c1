/* The C ABI from plain C (C11): compiles against include/pentago_amd.h alone, links
 * libpentago_amd.so, and calls the host-only entry points (no device needed). */
#include <stdio.h>
#include <string.h>

#include "pentago_amd.h"

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { fails++; printf("FAIL: " __VA_ARGS__); printf("\n"); } } while (0)

int main(void) {
  char digest[41];
  CHECK(pe_table_digest("rotations", digest) == PE_OK, "pe_table_digest: %s", pe_last_error());
  printf("rotations %s\n", digest);

  pe_section s33 = {{{3, 3}, {3, 3}, {3, 3}, {3, 3}}};
  int32_t shape[4];
  CHECK(pe_section_shape(&s33, shape) == PE_OK && shape[0] == 420 && shape[3] == 420, "section shape");

  pe_section root;
  memset(&root, 0, sizeof root);
  int n = 0;
  CHECK(pe_descendent_sections(&root, 35, 0, NULL, &n) == PE_OK && n == 330, "slice 35 has %d sections", n);
  pe_section secs[330];
  CHECK(pe_descendent_sections(&root, 35, 330, secs, &n) == PE_OK, "descendent sections");
  const int64_t bytes = pe_slice_bytes(n, secs);
  printf("slice 35: %d sections, %lld slab bytes\n", n, (long long)bytes);
  CHECK(bytes == 60354LL * 4096 * 64, "slice 35 slab bytes %lld (60354 blocks of 4096 nodes)", (long long)bytes);

  int order[330], ends[330], groups = 0;
  CHECK(pe_slice_plan_groups(n, secs, (int64_t)4 << 30, 0, order, ends, &groups) == PE_OK && groups > 1 &&
            ends[groups - 1] == n && order[0] == 0 && order[n - 1] == n - 1, "group planning");
  {
    /* the arena holds the largest group: at most the budget, at least its largest section */
    int64_t largest = 0;
    for (int k = 0; k < n; k++) {
      const int64_t b = pe_slice_bytes(1, &secs[k]);
      if (b > largest) largest = b;
    }
    const int64_t arena = pe_slice_arena_bytes(n, secs, (int64_t)4 << 30, 0);
    CHECK(arena >= largest && arena <= ((int64_t)4 << 30), "arena bytes %lld", (long long)arena);
    CHECK(pe_slice_arena_bytes(n, secs, 0, 0) == -1, "a zero budget is rejected");
  }

  pe_line line;
  memset(&line, 0, sizeof line);
  line.section = s33;
  line.dimension = 1;
  line.length = 53;
  pe_line_info info;
  CHECK(pe_line_describe(&line, &info) == PE_OK && info.output_nodes == 420LL * 8 * 8 * 8,
        "line describe: %s", pe_last_error());
  line.length = 7;
  CHECK(pe_line_describe(&line, &info) == PE_ERR_INVALID, "a wrong line length is rejected");
  printf("last error: %s\n", pe_last_error());

  /* an empty board with 18 stones of nobody is not a game state: rejected, not crashed */
  pe_high_board b = {{0, 0}, 0};
  int done = -1, value = 0;
  CHECK(pe_high_done(&b, &done, &value) == PE_OK && done == 0, "empty board is not done");
  printf("%s\n", fails ? "capi_check failed" : "capi_check ok");
  return fails != 0;
}
