"""Write the reference's test-vector arrays as JSON fixtures.

Reads the two ``include/tv.h`` headers of the reference checkout as plain text
(regular expressions over the C array initialisers; nothing is compiled,
imported or executed) and stores each array as a hex string:

* ``saber_tv.json`` from ``TMVP-Saber/SABER-GPU/include/tv.h``:
  ``A`` (tv.h:4, GenMatrix output, 2304 x uint16), ``sp`` (tv.h:6, GenSecret
  output, 768 x uint16), ``pk`` (tv.h:8, 992 B), ``m_tv`` (tv.h:9, 32 B),
  ``sk_tv`` (tv.h:10, 2304 B), ``buf_tv`` (tv.h:12, 64 B).
* ``sable_tv.json`` from ``TMVP_SABLE/include/tv.h``: ``buf_tv`` (tv.h:13),
  ``pk_tv`` (tv.h:15, 896 B), ``kr_tv`` (tv.h:17, 64 B), ``sk_tv`` (tv.h:19,
  1152 B), ``m_tv`` (tv.h:21, 32 B).

uint16 arrays are stored little-endian.  Usage::

    python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ARRAY_RE = re.compile(r"(uint8_t|uint16_t)\s+(\w+)\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}")


def parse_header(path):
    text = open(path, encoding="ascii", errors="strict").read()
    out = {}
    for m in ARRAY_RE.finditer(text):
        ctype, name, body = m.groups()
        vals = [int(v, 0) for v in body.replace("\n", " ").split(",") if v.strip()]
        if ctype == "uint16_t":
            raw = b"".join(v.to_bytes(2, "little") for v in vals)
        else:
            raw = bytes(vals)
        out[name] = {"ctype": ctype, "count": len(vals), "hex": raw.hex()}
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    jobs = {
        "saber_tv.json": os.path.join(ref, "TMVP-Saber/SABER-GPU/include/tv.h"),
        "sable_tv.json": os.path.join(ref, "TMVP_SABLE/include/tv.h"),
    }
    for fname, src in jobs.items():
        arrays = parse_header(src)
        doc = {"source": os.path.relpath(src, ref), "arrays": arrays}
        with open(os.path.join(HERE, fname), "w") as f:
            json.dump(doc, f, indent=1, sort_keys=True)
        print(fname, {k: v["count"] for k, v in arrays.items()})


if __name__ == "__main__":
    main()
