"""The host replay skips chunks whose inserts cannot change the heap
(search.cpp:job_replay).  The library's self-check replays random per-pair
tier results both ways -- with the chunk filter and with every chunk's lane
schedules simulated in full -- and counts the cases whose hit lists or
overflow counts differ.  The full simulation is the one the GPU parity tests
pin against the oracle; here the filter must reproduce it exactly, including
tie order at the k-th boundary (narrow score ranges), the cascade's duplicate
inserts (several query buffers with 8- and 16-bit overflows) and k up to 5000.
Host only: no GPU call."""
import pytest

ssa = pytest.importorskip("libssa_amd")


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_chunk_filter_matches_full_replay(seed):
    assert ssa.selfcheck_replay(seed, 500) == 0
