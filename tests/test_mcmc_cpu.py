"""CPU pins of the MCMC reference used by the GPU chain tests."""
import numpy as np

from mcmc_ref import normals, philox4x32_10


def test_philox_known_answers():
    # Known-answer vectors published with the Random123 library (kat_vectors)
    assert philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_normals_are_standard():
    z = np.concatenate([normals(s, 3, 64, 7) for s in range(400)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
