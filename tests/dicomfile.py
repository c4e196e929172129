"""Minimal reader for the DICOM files ctbb_write_dicom_series writes
(Part-10, explicit VR little endian), for the tests: no DICOM library is
installed, so the checks parse the byte stream themselves."""
import re
import struct

import numpy as np

LONG_VRS = ("OB", "OW", "UN", "SQ", "UT")


def read(path):
    """-> (meta elements, data set elements, tag order); elements map
    (group, element) -> (VR, value bytes)."""
    b = open(path, "rb").read()
    assert b[:128] == bytes(128), "preamble must be 128 zero bytes"
    assert b[128:132] == b"DICM"
    pos = 132
    meta, data, order = {}, {}, []
    while pos < len(b):
        g, e = struct.unpack_from("<HH", b, pos)
        vr = b[pos + 4:pos + 6].decode("ascii")
        if vr in LONG_VRS:
            assert b[pos + 6:pos + 8] == b"\0\0"
            n = struct.unpack_from("<I", b, pos + 8)[0]
            pos += 12
        else:
            n = struct.unpack_from("<H", b, pos + 6)[0]
            pos += 8
        assert n % 2 == 0, f"odd value length in ({g:04x},{e:04x})"
        val = b[pos:pos + n]
        assert len(val) == n, "truncated element"
        pos += n
        (meta if g == 2 else data)[(g, e)] = (vr, val)
        order.append((g, e))
    return meta, data, order


def text(el):
    vr, v = el
    return v.rstrip(b"\0" if vr == "UI" else b" ").decode("ascii")


def ds(el):
    vals = text(el).split("\\")
    assert all(len(x) <= 16 for x in vals), "DS value longer than 16 characters"
    return [float(x) for x in vals]


def us(el):
    assert el[0] == "US"
    return struct.unpack("<H", el[1])[0]


def pixels(el, rows, cols):
    assert el[0] == "OW"
    return np.frombuffer(el[1], dtype="<i2").reshape(rows, cols)


UID_RE = re.compile(r"^(0|[1-9][0-9]*)(\.(0|[1-9][0-9]*))*$")


def valid_uid(u):
    return len(u) <= 64 and UID_RE.match(u) is not None
