// Work-queue decode (f32) for launches whose codes are BG1 codes of Z <= 256.
#include "ldpc_persist.h"

namespace syn {

const void *persist_kernel_f32_small() { return (const void *)persist_decode<false, 1>; }

}  // namespace syn
