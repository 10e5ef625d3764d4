"""The router's visited cache codes (router_device.hpp vis_key) identify a
vertex exactly: for every vid below 2^B, (slot, code) of choice 1 and of
choice 2 are each injective, and the two choices' code ranges do not overlap,
so a lookup can only hit on the vertex that was stored.  This is the property
that keeps search results independent of the cache (a false hit would skip a
vertex the reference visits)."""
import numpy as np
import pytest


def codes(v, B, mode):
    # Same arithmetic as vis_key in artea_amd/csrc/router_device.hpp.
    hm = np.uint32((1 << B) - 1)
    h1 = (v * np.uint32(0x9E3779B1)) & hm
    h2 = (v * np.uint32(0x85EBCA77)) & hm
    if mode == 0:
        q1, q2 = h1 // 7680, h2 // 7680
        return h1 - q1 * 7680, q1 + 1, h2 - q2 * 7680, np.where(q2 <= 117, q2 + 138, 0)
    q1, q2 = h1 // 3840, h2 // 3840
    return h1 - q1 * 3840, q1 + 1, h2 - q2 * 3840, np.where((mode == 1) & (q2 <= 30581), q2 + 34954, 0)


@pytest.mark.parametrize("B,mode", [(13, 0), (16, 0), (20, 0), (21, 1), (24, 1)])
def test_visited_codes_identify_the_vertex(B, mode):
    v = np.arange(1 << B, dtype=np.uint32)
    s1, t1, s2, t2 = codes(v, B, mode)
    width, slots = (256, 7680) if mode == 0 else (65536, 3840)
    assert s1.max() < slots and s2.max() < slots
    assert t1.min() >= 1 and t1.max() < width
    assert len(np.unique(s1.astype(np.uint64) * width + t1)) == len(v)
    second = t2 != 0
    assert second.mean() > 0.85  # most vertices have a second choice
    assert t2[second].max() < width and t2[second].min() > t1.max()
    assert len(np.unique(s2[second].astype(np.uint64) * width + t2[second])) == second.sum()


def test_visited_code_bounds_at_the_mode_limits():
    # The largest quotients of each mode still fit their code ranges.
    assert (2**20 - 1) // 7680 + 1 == 137 and 117 + 138 == 255
    assert (2**27 - 1) // 3840 + 1 == 34953 and 30581 + 34954 == 65535
