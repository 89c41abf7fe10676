"""Pure-Python restatement of the reference's FASTA reader and dedup.

TEST INFRASTRUCTURE ONLY (the checker for small inputs): only tests/ may
import it; the product never does.

read_fasta       src/hamming_impl.cc:200-242
fasta_indices    src/hamming.cc:76-102
"""
from __future__ import annotations


class _Stream:
    """std::getline over a byte string with the stream's eof flag: a last
    line without '\\n' succeeds and sets eof; a read at the end fails."""

    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0
        self.eof = False

    def getline(self):
        if self.pos >= len(self.data):
            self.eof = True
            return None
        nl = self.data.find(b"\n", self.pos)
        if nl < 0:
            line = self.data[self.pos:]
            self.pos = len(self.data)
            self.eof = True
            return line
        line = self.data[self.pos:nl]
        self.pos = nl + 1
        return line


def _records(data: bytes, n: int):
    s = _Stream(data)
    s.getline()  # skip the first line
    if n == 0:
        n = float("inf")
    count = 0
    while count < n and not s.eof:
        seq = []
        while True:
            line = s.getline()
            # line[0] of an empty std::string is '\0', never '>'
            if line is None or (len(line) > 0 and line[0:1] == b">"):
                break
            seq.append(line)
        yield b"".join(seq).decode("latin-1")
        count += 1


def read_fasta(path: str, remove_duplicates: bool = False, n: int = 0):
    """(data, sequence_indices) as read_fasta returns them."""
    with open(path, "rb") as f:
        raw = f.read()
    return read_fasta_bytes(raw, remove_duplicates, n)


def read_fasta_bytes(raw: bytes, remove_duplicates: bool = False, n: int = 0):
    """read_fasta on the bytes of a file."""
    data, indices, index = [], [], {}
    for seq in _records(raw, n):
        if remove_duplicates:
            if seq not in index:
                index[seq] = len(data)
                data.append(seq)
            indices.append(index[seq])
        else:
            data.append(seq)
    return data, indices


def fasta_indices(path: str, n: int = 0):
    return read_fasta(path, True, n)[1]
