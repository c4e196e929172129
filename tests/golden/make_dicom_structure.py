"""Writes tests/golden/dicom/ct_image_structure.json: the structure of the
CT Image Storage file the reference ships (resources/testfile.dcm): its
file-meta and data-set tags with their VRs in file order, the transfer
syntax and SOP class UIDs and the Image Pixel module values.  No patient,
study or pixel values are kept.  Run where the reference tree exists:

    python tests/golden/make_dicom_structure.py [/root/reference]
"""
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import dicomfile as D  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    meta, data, order = D.read(os.path.join(ref, "resources", "testfile.dcm"))
    tag = lambda g, e: f"{g:04x},{e:04x}"
    pixel = {}
    for name, (g, e) in {"samples_per_pixel": (0x28, 0x2), "bits_allocated": (0x28, 0x100),
                         "bits_stored": (0x28, 0x101), "high_bit": (0x28, 0x102),
                         "pixel_representation": (0x28, 0x103)}.items():
        pixel[name] = struct.unpack("<H", data[(g, e)][1])[0]
    pixel["photometric_interpretation"] = D.text(data[(0x28, 0x4)])
    out = {
        "source": "reference resources/testfile.dcm (structure only)",
        "transfer_syntax": D.text(meta[(0x2, 0x10)]),
        "sop_class": D.text(data[(0x8, 0x16)]),
        "media_storage_sop_class": D.text(meta[(0x2, 0x2)]),
        "meta": [[tag(g, e), meta[(g, e)][0]] for (g, e) in order if g == 2],
        "dataset": [[tag(g, e), data[(g, e)][0]] for (g, e) in order if g != 2],
        "pixel_module": pixel,
    }
    path = os.path.join(HERE, "dicom", "ct_image_structure.json")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
