// Constants shared by the host setup (g++) and the device code (hipcc).
#pragma once

namespace ugmg {
constexpr int kSlice = 64;   // rows per SELL slice = one wavefront
constexpr int kBlock = 256;  // threads per block of the row kernels
}  // namespace ugmg
