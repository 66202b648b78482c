"""The autotuner's variant policy, on made-up timings (CPU only).

Systems of different kernel variants in one batch run in separate launches,
so each mode takes one uniform variant (the smallest sum of times normalised
by each bucket's best), a bucket keeps its own best only where that is at
least OVERRIDE_GAIN faster, and the buckets the tuner does not time take the
uniform variant.
"""
import pytest

from nbody_amd import autotune


def test_uniform_variant_minimises_normalised_sum():
    # variant 1 is best in two of three buckets; variant 2 wins one bucket by
    # a little but loses badly elsewhere
    times = {6: [1.00, 0.90, 0.95],
             7: [2.00, 1.80, 3.00],
             8: [4.00, 4.10, 3.95]}
    uniform, pick = autotune.choose_variants(times)
    assert uniform == 1
    # bucket 8: variant 2 is only 3.8 % faster than the uniform one
    assert pick == {6: 1, 7: 1, 8: 1}


def test_bucket_override_needs_the_gain():
    gain = autotune.OVERRIDE_GAIN
    base = 1.0
    just_below = base / (1.0 + gain) * 1.001   # a bit less than the gain
    clearly = base / (1.0 + gain) * 0.99       # a bit more than the gain
    times = {6: [base, 2.0], 7: [base, 2.0], 8: [base, just_below], 9: [base, clearly]}
    uniform, pick = autotune.choose_variants(times)
    assert uniform == 0
    assert pick[8] == 0 and pick[9] == 1


def test_untimed_buckets_take_the_uniform_variant():
    table = autotune.variant_table(4, {6: 4, 7: 5, 20: 4}, 32)
    assert len(table) == 32
    assert table[7] == 5
    assert all(table[b] == 4 for b in range(32) if b != 7)


@pytest.mark.parametrize("uniform", [0, 2, 8])
def test_variant_table_keeps_every_pick(uniform):
    pick = {b: (b * 7) % 9 for b in range(6, 21)}
    table = autotune.variant_table(uniform, pick, 32)
    for b in range(32):
        assert table[b] == pick.get(b, uniform)
