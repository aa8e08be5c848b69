"""The generic tiles find the exponent block of tile value i as
umulhi(i, ceil(2^32 / B)) (frsz_amd/csrc/gtile.hpp, block_of): exact for every
value index a tile holds (i < 4096) and every block size it takes
(2 <= B <= 4096), checked exhaustively."""
import numpy as np


def test_multiply_high_block_index_is_exact():
    i = np.arange(4096, dtype=np.uint64)
    for B in range(2, 4097):
        magic = np.uint64(-(-(1 << 32) // B))
        assert np.array_equal((i * magic) >> np.uint64(32), i // np.uint64(B)), B
