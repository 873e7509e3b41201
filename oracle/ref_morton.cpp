// ref_morton.cpp -- test infrastructure: the Morton encoding the reference's
// keys are built from (morton_id<Depth, Rank>::encode, include/morton/
// morton_id.hpp:131-139 and 349-358: libmorton's 64-bit encode of the
// coordinates at the 2^Depth resolution, shifted left by 6, OR the level),
// from the reference's vendored libmorton headers where they lie
// (lib/libmorton; built by `make -C oracle ref` into oracle/_ref/, which is
// not committed and is opened only by the CPU tests of
// tests/test_oracle_ref.py).  morton_id.hpp itself needs C++23 library functions this image
// lacks, so the key is assembled here from libmorton exactly as it writes it.
#include <libmorton/morton.h>

#include <cstdint>

extern "C"
{

uint64_t ref_morton_key(int rank, uint32_t x, uint32_t y, uint32_t z, uint32_t level)
{
    const uint64_t m = rank == 2 ? libmorton::morton2D_64_encode(x, y) : libmorton::morton3D_64_encode(x, y, z);
    return (m << 6) | (level & 0x3Fu);
}

} // extern "C"
