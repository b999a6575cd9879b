"""Device functors compile with the chain's kernels on the CPU (mci_check_device_functor: hiprtc
for gfx950, no GPU): the fixtures' sampling function, observable and trial moves build, a broken
one reports the compiler's message, and names must be C identifiers."""
import ctypes as C

import device_functors as DF
from mcitest import load

OK = 0


def check(name, src, what):
    L = load("mci")
    L.mci_check_device_functor.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int64]
    L.mci_check_device_functor.restype = C.c_int
    buf = C.create_string_buffer(1 << 16)
    st = L.mci_check_device_functor(name.encode(), src.encode(), int(what), buf, len(buf))
    return st, buf.value.decode(), L.mci_last_error().decode()


def test_fixture_functors_compile():
    for name, src, what in (("quartic", DF.QUARTIC, 0), ("moments", DF.MOMENTS, 1), ("pair", DF.PAIR, 2), ("scale", DF.SCALE, 2)):
        st, log, _ = check(name, src, what)
        assert st == OK, (name, log)


def test_broken_functor_reports_the_compiler():
    st, log, err = check("broken", "__device__ void broken_observable(int n, MciVec x, const double *p, MciVec o) { o[0] = y; }", 1)
    assert st != OK
    assert "undeclared identifier 'y'" in log and "undeclared identifier 'y'" in err


def test_functor_names_are_identifiers():
    for bad in ("", "1abc", "a-b", "x y"):
        st, _, err = check(bad, DF.MOMENTS, 1)
        assert st != OK and "C identifier" in err
