"""The committed fixtures hold data only.

A reference state file carries the reference's parameter file, build
configuration and source diffs in its HDF5 userblock.  The fixture script
zeroes that block; this test keeps any fixture under tests/golden from
carrying such text again (raw or inside an .npz member)."""
import os
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

MARKERS = (b"{[( START USERBLOCK", b"{[( COMPRESSED", b"{[( INIFILE", b"diff --git", b"SUBROUTINE")


def _blobs():
    for name in sorted(os.listdir(GOLDEN)):
        path = os.path.join(GOLDEN, name)
        if not os.path.isfile(path) or name.endswith(".py"):
            continue
        with open(path, "rb") as f:
            yield name, f.read()
        if name.endswith(".npz"):
            with zipfile.ZipFile(path) as z:
                for member in z.namelist():
                    yield name + ":" + member, z.read(member)


def test_no_reference_text_in_fixtures():
    seen = 0
    for name, blob in _blobs():
        seen += 1
        for m in MARKERS:
            assert m not in blob, "%s holds %r" % (name, m)
    assert seen >= 2


def test_state_fixture_userblock_is_zero():
    path = os.path.join(GOLDEN, "2Dplasma_test_State_000.00000005000000000.h5")
    with open(path, "rb") as f:
        blob = f.read()
    off = blob.find(b"\x89HDF\r\n\x1a\n")
    assert off in (0, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536)
    assert blob[:off] == bytes(off)
