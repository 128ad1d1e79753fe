// Work-queue decode (int8 messages, i16 posterior).
#include "ldpc_persist.h"

namespace syn {

const void *persist_kernel_i8() { return (const void *)persist_decode<true, 0>; }

}  // namespace syn
