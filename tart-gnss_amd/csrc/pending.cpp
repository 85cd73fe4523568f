// Entry points whose GPU implementation lands in the following commits.
#include "internal.h"

extern "C" {
int tga_rfi(const uint8_t*, size_t, size_t, int64_t, int32_t, double, tga_rfi_channel*) {
  return TGA_ERR_UNSUPPORTED;
}
int tga_acquire_all_json(tga_plan*, const uint8_t*, size_t, size_t, const double*,
                         const int64_t*, int32_t, char**) {
  return TGA_ERR_UNSUPPORTED;
}
void tga_free_string(char* s) { delete[] s; }
}
