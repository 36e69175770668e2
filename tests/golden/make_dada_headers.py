"""Copies the DADA header blocks (first 4096 bytes) of the reference's
MultiFileReader fixtures (cpp/skyweaver/test/data/dada_files/data_chunk_*.dada,
used by cpp/skyweaver/test/src/MultiFileReaderTester.cu:10-19) into
tests/golden/dada/.  The voltages after the headers are not kept: the reader
tests write their own payloads behind these headers."""
import os
import sys

SRC = "/root/reference/cpp/skyweaver/test/data/dada_files"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dada")


def main():
    os.makedirs(DST, exist_ok=True)
    names = sorted(n for n in os.listdir(SRC) if n.startswith("data_chunk_") and n.endswith(".dada"))
    for n in names:
        with open(os.path.join(SRC, n), "rb") as f:
            hdr = f.read(4096)
        size = os.path.getsize(os.path.join(SRC, n))
        out = os.path.join(DST, n.replace(".dada", ".hdr"))
        with open(out, "wb") as f:
            f.write(hdr)
        print(out, "data bytes in the fixture:", size - 4096)
    return 0


if __name__ == "__main__":
    sys.exit(main())
