"""The reference's transfer-cache byte layout (TRF1), for transfer sets built
on the device.

``Rad/Bounce/TransferSetCodec.cs`` stores a ``TransferSet`` as one index blob
and a list of chunks:

* index (``Index``, :30-42): the magic ``"TRF1"`` (``0x31465254``), the patch
  count, the longest list (``Max``), the total (int64), then every patch's
  count, all little-endian;
* chunks (``Chunks``, :49-98): every transfer, patch after patch, as 8 bytes
  (int32 source patch, float32 weight bits), cut into chunks of at most
  ``max_chunk_bytes`` (rounded down to whole transfers, at least one);
* packed chunks (``Pack``/``Unpack``, :107-185): the transfer count (int32),
  then a zlib stream of the chunk with the patches delta-coded and the four
  bytes of every delta and every weight grouped by significance.

The records themselves come from the device (``vh_transfers_export``); this
module does the host-side framing, and reads sets back with the reference's
checks (``Read``, :369-445): a set that does not fit its index is refused.
"""
from __future__ import annotations

import struct
import zlib
from typing import List, Optional, Sequence, Tuple

import numpy as np

MAGIC = 0x31465254  # "TRF1"
TRANSFER_BYTES = 8
_INT_MAX = 2**31 - 1
_ARRAY_MAX = 0x7FFFFFC7  # .NET Array.MaxLength


def index(counts: np.ndarray) -> bytes:
    """The index blob for per-patch list lengths ``counts``."""
    counts = np.ascontiguousarray(counts, dtype="<i4")
    total = int(counts.astype(np.int64).sum())
    longest = int(counts.max()) if counts.size else 0
    return struct.pack("<Iiiq", MAGIC, counts.size, longest, total) + counts.tobytes()


def chunks(records: bytes, max_chunk_bytes: int) -> List[bytes]:
    """``records`` (8 bytes a transfer, patch after patch) cut as the
    reference cuts them: chunks of ``clamp(max_chunk_bytes / 8, 1, ...)``
    transfers, the last one shorter."""
    mv = memoryview(records)
    if len(mv) % TRANSFER_BYTES:
        raise ValueError("records are 8 bytes each")
    per = int(min(max(max_chunk_bytes // TRANSFER_BYTES, 1), _INT_MAX // TRANSFER_BYTES))
    step = per * TRANSFER_BYTES
    return [bytes(mv[k:k + step]) for k in range(0, len(mv), step)]


def pack(chunk: bytes) -> bytes:
    """A raw chunk packed for storage (``TransferSetCodec.Pack``)."""
    rec = np.frombuffer(chunk, dtype="<u4").reshape(-1, 2)
    n = rec.shape[0]
    patch = rec[:, 0].view("<i4").astype(np.int64)
    prev = np.concatenate([[0], patch[:-1]])
    delta = ((patch - prev) & 0xFFFFFFFF).astype("<u4")
    planes = np.empty((8, n), np.uint8)
    for b in range(4):
        planes[b] = (delta >> (8 * b)) & 0xFF
        planes[4 + b] = (rec[:, 1] >> (8 * b)) & 0xFF
    return struct.pack("<i", n) + zlib.compress(planes.tobytes(), 1)


def unpack(packed: bytes) -> Optional[bytes]:
    """``TransferSetCodec.Unpack``: the raw chunk, or None when the bytes are
    not a packed chunk (too short, a bad count, a stream that does not
    inflate to exactly the count, or trailing bytes)."""
    if len(packed) < 4:
        return None
    (n,) = struct.unpack_from("<i", packed)
    if n < 0 or n > _ARRAY_MAX // TRANSFER_BYTES:
        return None
    d = zlib.decompressobj()
    try:
        raw = d.decompress(packed[4:])
        raw += d.flush()
    except zlib.error:
        return None
    if not d.eof or d.unused_data or len(raw) != n * TRANSFER_BYTES:
        return None
    planes = np.frombuffer(raw, np.uint8).reshape(8, n).astype(np.uint32)
    delta = np.zeros(n, np.uint32)
    weight = np.zeros(n, np.uint32)
    for b in range(4):
        delta |= planes[b] << (8 * b)
        weight |= planes[4 + b] << (8 * b)
    patch = (np.cumsum(delta.astype(np.int64)) & 0xFFFFFFFF).astype(np.uint32)
    out = np.empty((n, 2), "<u4")
    out[:, 0] = patch
    out[:, 1] = weight
    return out.tobytes()


def read(index_blob: bytes, chunk_list: Sequence[bytes], patch_count: int
         ) -> Optional[Tuple[np.ndarray, np.ndarray, np.ndarray, int]]:
    """``TransferSetCodec.Read``: (counts, source patches, weights, max), or
    None when the index and chunks do not fit together: a wrong magic or
    patch count, a chunk that is not whole transfers, bytes that do not add
    up to the total, a negative count, counts that do not add up to the total
    or whose longest is not ``max``, or a patch out of range."""
    if (len(index_blob) < 20 or struct.unpack_from("<I", index_blob)[0] != MAGIC
            or struct.unpack_from("<i", index_blob, 4)[0] != patch_count
            or len(index_blob) != 20 + 4 * patch_count):
        return None
    longest_stated = struct.unpack_from("<i", index_blob, 8)[0]
    total = struct.unpack_from("<q", index_blob, 12)[0]
    nbytes = 0
    for c in chunk_list:
        if len(c) % TRANSFER_BYTES:
            return None
        nbytes += len(c)
    if total < 0 or nbytes != total * TRANSFER_BYTES or total > _ARRAY_MAX:
        return None
    counts = np.frombuffer(index_blob, "<i4", count=patch_count, offset=20).astype(np.int32)
    if (counts < 0).any():
        return None
    if int(counts.astype(np.int64).sum()) != total or (int(counts.max()) if patch_count else 0) != longest_stated:
        return None
    rec = np.frombuffer(b"".join(chunk_list), "<u4").reshape(-1, 2)
    src = rec[:, 0].view("<i4").astype(np.int32)
    if ((src < 0) | (src >= patch_count)).any():
        return None
    return counts, src, rec[:, 1].copy().view("<f4").astype(np.float32), longest_stated
