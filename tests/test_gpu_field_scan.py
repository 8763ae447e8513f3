"""GPU field scan (molscat_amd.driver.field_scan) on the two N = 48 field
decks (one constant term, the field; decks.field48 with its L structure
and decks.hyperfine48, the s-wave deck with random orthogonal singlet and
triplet projectors): a(B) at several fields in one batched call against the
oracle's S-matrix of each field turned into CALCA's scattering length, and
the batch of 4096 fields (the scan's full size) finishing with every
instance's status 0, a(B) identical to a small batch's at shared fields (on
the same MDPROP kernel) and equal to the oracle's at three fields inside the
batch."""
import numpy as np
import pytest

from molscat_amd import decks
from molscat_amd import driver as D
from molscat_amd import system as S
from oracle import molscat_oracle as O

from test_oracle_ytrans import _oracle_block

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from molscat_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("deck", ["field48", "hyperfine48"])
def test_scattering_length_vs_oracle(eng, deck):
    fb = getattr(decks, deck)()
    fields = [0.0, 137.0, 512.0, 1000.0] if deck == "field48" else [0.25, 137.0, 512.0, 1000.0]
    print(f"field scan N={fb.n} at {fields} G", flush=True)
    fs = D.field_scan(eng, fb, fields, 3.5, 25.0, 0.02)
    assert np.all(fs.status == 0)
    nstep, dr = S.drset_constant(3.5, 25.0, 0.02)
    for m, f in enumerate(fields):
        ob = _oracle_block(fb, f)
        ref = O.smatrix_ld(ob, fs.ered[m], O.Grid(3.5, 25.0, nstep, dr), iprop=6)
        assert int(ref["NOPEN"]) == int(fs.nopen[m])
        e0 = int(ref["NB"][0])
        assert int(fb.lvals[e0]) == 0
        aref = O.calca(ref["SR"][0, 0], ref["SI"][0, 0], ref["WVEC"][e0], 0)
        d = abs(fs.a[m] - aref) / abs(aref)
        print(f"  B={f} G a={fs.a[m].real:.6f}{fs.a[m].imag:+.2e}i (oracle {aref.real:.6f}) rel {d:.1e}",
              flush=True)
        assert d < 1e-8


@pytest.mark.parametrize("deck", ["field48", "hyperfine48"])
def test_full_scan_4096_fields(eng, deck):
    fb = getattr(decks, deck)()
    fields = np.linspace(0.0, 1000.0, 4096)
    print(f"field scan ({deck}): 4096 fields in one call", flush=True)
    fs = D.field_scan(eng, fb, fields, 3.5, 25.0, 0.02)
    assert np.all(fs.status == 0) and np.all(np.isfinite(fs.a))
    # a batch of three on the kernel the 4096-field batch ran on (the
    # register-resident MDPROP, chosen automatically from 512 instances)
    # gives the same a(B) bitwise
    eng.set_md_kernel(2)
    try:
        sub = D.field_scan(eng, fb, fields[[0, 1000, 4095]], 3.5, 25.0, 0.02)
    finally:
        eng.set_md_kernel(0)
    assert np.array_equal(sub.a, fs.a[[0, 1000, 4095]])
    # members of the full batch against the oracle
    nstep, dr = S.drset_constant(3.5, 25.0, 0.02)
    for m in (1, 2047, 4095):
        ob = _oracle_block(fb, fields[m])
        ref = O.smatrix_ld(ob, fs.ered[m], O.Grid(3.5, 25.0, nstep, dr), iprop=6)
        e0 = int(ref["NB"][0])
        aref = O.calca(ref["SR"][0, 0], ref["SI"][0, 0], ref["WVEC"][e0], 0)
        d = abs(fs.a[m] - aref) / abs(aref)
        print(f"  [{m}] B={fields[m]:.4f} G a={fs.a[m].real:.6f} (oracle {aref.real:.6f}) rel {d:.1e}", flush=True)
        assert d < 1e-8
