"""Host-side configuration logic (no GPU): the seeded generator and the ensemble draw."""
import mpdata_amd.configs as K


def test_splitmix64_known_answer():
    # splitmix64 reference outputs for seed 1234567 (Steele, Lea & Flood 2014; the
    # published C reference implementation by S. Vigna)
    r = K.SplitMix64(1234567)
    assert [r.next() for _ in range(3)] == [6457827717110365317, 3203168211198807973,
                                            9817491932198370423]


def test_ensemble_draw_is_seeded_and_uses_the_five_shapes():
    a = K.ensemble_dims(192, 4)
    assert a == K.ensemble_dims(192, 4)
    assert a != K.ensemble_dims(192, 5)
    assert set(a) <= set(K.ENSEMBLE_SHAPES) and len(set(a)) == 5


def test_random_modes_are_bounded_and_nonzero():
    r = K.SplitMix64(3)
    modes = K.random_modes(r, 4, 1.0)
    assert len(modes) == 3 and all(len(m) == 4 for m in modes)
    for comp in modes:
        for kx, ky, kz, amp, ph in comp:
            assert (kx, ky, kz) != (0, 0, 0) and max(abs(kx), abs(ky), abs(kz)) <= 3
            assert 0 < amp <= 1.0 and 0 <= ph < 6.3
