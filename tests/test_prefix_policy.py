"""Host logic of the prefix-list policy of the drop-in API (no GPU): the
divisor each frame uses given the outcomes of earlier frames."""
from fastergs_amd.rasterizer import _PrefixPolicy


def _run(pol, fails_above=None, fails_always=False, frames=40):
    """Frames whose outcome is known before the next one starts; a divisor
    fails when it is above fails_above (or always). Returns the divisors used."""
    used = []
    for _ in range(frames):
        d = pol.divisor()
        used.append(d)
        if d > 1:
            pol._outcome(d, fails_always or (fails_above is not None and d > fails_above))
    return used


def test_ramps_up_to_the_largest_divisor_that_suffices():
    pol = _PrefixPolicy(16)
    used = _run(pol, fails_above=64)
    # 4 frames at 16, a try at 32, 4 at 32, a try at 64, then a failed try at 128
    assert used[:5] == [16, 16, 16, 16, 32], used
    assert 128 in used and used.count(128) == 1, used
    assert used[-10:] == [64] * 10, used
    assert pol.good == 64 and pol.cap == 64


def test_never_above_the_cap():
    pol = _PrefixPolicy(16)
    used = _run(pol, frames=200)
    assert max(used) == _PrefixPolicy.MAX_DIV and used[-1] == _PrefixPolicy.MAX_DIV


def test_shape_without_a_sufficing_prefix_turns_off():
    pol = _PrefixPolicy(16)
    used = _run(pol, fails_always=True, frames=10)
    assert used[0] == 16 and used[1:] == [0] * 9, used
    assert pol.off == _PrefixPolicy.RETRY - 9


def test_retry_after_the_whole_list_frames():
    pol = _PrefixPolicy(16)
    used = _run(pol, fails_always=True, frames=_PrefixPolicy.RETRY + 3)
    assert used[_PrefixPolicy.RETRY + 1] == 16, used[-4:]


def test_kept_divisor_that_stops_sufficing_is_halved():
    pol = _PrefixPolicy(16)
    _run(pol, fails_above=32, frames=30)
    assert pol.good == 32
    pol._outcome(32, True)  # the scene changed: 32 no longer suffices
    assert pol.good == 16 and pol.cap == 16 and pol.divisor() == 16
