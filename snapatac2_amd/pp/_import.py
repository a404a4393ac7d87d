"""Drop-in for snapatac2.pp.import_fragments, parsed and grouped on the GPU.

Reference: snapatac2-python/python/snapatac2/preprocessing/_basic.py:156-340
(Python API), snapatac2-python/src/preprocessing.rs:87-183 (binding,
shift_fragment), snapatac2-core/src/preprocessing/import.rs:31-223 (grouping,
QC, CSR layout). The fragment text is read (and gunzipped) on the host, copied
to the device once, and libsnapgpu does the rest: line index, field parsing,
chromosome lookup, grouping by barcode, QC sums, filters and the CSR fill.
"""
from __future__ import annotations

import ctypes as C
import gzip
import logging
from pathlib import Path

import numpy as np

from .. import _lib
from .._device import DeviceFragments
from ..anndata import AnnData
from ..genome import reference_sequences, sorted_chrom_sizes

FRAGMENT_SINGLE = "fragment_single"  # snapatac2-core/src/feature_count/mod.rs:20
FRAGMENT_PAIRED = "fragment_paired"  # mod.rs:23

log = logging.getLogger(__name__)


def _name_table(names, lengths=None, mito=None):
    """Host-packed hashed name table (snapgpu_frag_table_pack), as a uint8 array."""
    L = _lib.lib()
    enc = [n.encode() if isinstance(n, str) else bytes(n) for n in names]
    arr = (C.c_char_p * max(len(enc), 1))(*enc)
    nbytes = L.snapgpu_frag_table_bytes(arr, len(enc))
    if nbytes < 0:
        raise ValueError("bad name table")
    out = np.zeros(nbytes, dtype=np.uint8)
    ln = None if lengths is None else np.ascontiguousarray(lengths, dtype=np.uint64)
    mt = None if mito is None else np.ascontiguousarray(mito, dtype=np.uint8)
    _lib.check(L.snapgpu_frag_table_pack(arr, None if ln is None else ln.ctypes.data,
                                         None if mt is None else mt.ctypes.data, len(enc),
                                         1 if lengths is not None else 0, out.ctypes.data, nbytes))
    return out


def _to_device_text(text, torch):
    """uint8 device tensor holding the text, 16-byte aligned, length padded to 16."""
    dev = torch.device("cuda")
    if isinstance(text, torch.Tensor):
        if text.device.type != "cuda" or text.dtype != torch.uint8 or text.dim() != 1:
            raise ValueError("text tensor must be a 1-D uint8 CUDA tensor")
        if text.data_ptr() % 16 == 0:
            return text, text.numel()
        out = torch.empty((text.numel() + 15) // 16 * 16, dtype=torch.uint8, device=dev)
        out[:text.numel()].copy_(text)
        return out, text.numel()
    buf = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray, memoryview)) \
        else np.ascontiguousarray(text, dtype=np.uint8).ravel()
    n = buf.size
    host = torch.empty((n + 15) // 16 * 16, dtype=torch.uint8, pin_memory=True)
    if n:
        host.numpy()[:n] = buf
    return host.to(dev, non_blocking=True), n


def import_fragment_text(text, chrom_sizes: dict[str, int], *, min_num_fragments: int = 200,
                         sorted_by_barcode: bool = True, whitelist=None,
                         chrM=("chrM", "M"), shift_left: int = 0, shift_right: int = 0,
                         n_bytes: int | None = None, return_summary: bool = False):
    """Fragment-file text (bytes, numpy uint8 or a uint8 CUDA tensor, of which
    the first n_bytes are used if given) -> AnnData with
    obsm['fragment_paired'|'fragment_single'] on the device, obs n_fragment /
    frac_dup / frac_mito and uns['reference_sequences']."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    dev = torch.device("cuda")
    cs = sorted_chrom_sizes(chrom_sizes)
    mito = set(chrM)
    chroms = _name_table(list(cs), list(cs.values()), [1 if c in mito else 0 for c in cs])
    d_chroms = torch.from_numpy(chroms).to(dev)
    d_wl = None
    if whitelist is not None:
        d_wl = torch.from_numpy(_name_table(sorted(set(whitelist)))).to(dev)
    d_text, size = _to_device_text(text, torch)
    n_bytes = size if n_bytes is None else int(n_bytes)
    if not 0 <= n_bytes <= size:
        raise ValueError("n_bytes exceeds the text")

    n_lines = C.c_int64(0)
    cws = torch.empty(L.snapgpu_frag_count_workspace_bytes(n_bytes), dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_frag_count_lines(d_text.data_ptr(), n_bytes, cws.data_ptr(), cws.numel(),
                                          C.byref(n_lines), st))
    del cws
    wsb = L.snapgpu_frag_workspace_bytes(n_bytes, n_lines.value)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    opt = _lib.FragOptions(int(shift_left), int(shift_right), max(int(min_num_fragments), 0),
                           1 if sorted_by_barcode else 0, 0)
    summ = _lib.FragSummary()
    _lib.check(L.snapgpu_frag_index(d_text.data_ptr(), n_bytes, n_lines.value, d_chroms.data_ptr(),
                                    None if d_wl is None else d_wl.data_ptr(), C.byref(opt),
                                    ws.data_ptr(), wsb, C.byref(summ), st))
    R, nnz, width = summ.n_rows, summ.nnz, max(summ.name_width, 1)
    cell_ptr = torch.empty(R + 1, dtype=torch.int64, device=dev)
    pos = torch.empty(nnz, dtype=torch.int32, device=dev)
    val = torch.empty(nnz, dtype=torch.int32, device=dev)
    qc = torch.empty((R, 3), dtype=torch.int64, device=dev)
    names = torch.empty((R, width), dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_frag_fill(d_text.data_ptr(), n_bytes, ws.data_ptr(), wsb, C.byref(summ),
                                   cell_ptr.data_ptr(), pos.data_ptr() if nnz else None,
                                   val.data_ptr() if nnz else None, qc.data_ptr() if R else None,
                                   names.data_ptr() if R else None, width, st))
    del ws
    adata = _to_anndata(cell_ptr, pos, val, qc, names, summ, cs, torch)
    return (adata, summ) if return_summary else adata


def _to_anndata(cell_ptr, pos, val, qc, names, summ, cs, torch):
    import pandas as pd
    R = summ.n_rows
    if R == 0:
        log.warning("No barcodes passed the QC filter. No data is imported.")  # import.rs:109-111
        return AnnData(n_obs=0, n_vars=0)
    q = qc.cpu().numpy().view(np.uint64)
    u, t, m = q[:, 0], q[:, 1], q[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        # FragmentQCBuilder::finish (qc.rs:452-465)
        frac_dup = 1.0 - (u + m).astype(np.float64) / t.astype(np.float64)
        frac_mito = m.astype(np.float64) / (u + m).astype(np.float64)
    nm = names.cpu().numpy()
    bcs = [bytes(r).rstrip(b"\0").decode() for r in nm]
    obs = pd.DataFrame({"n_fragment": u.copy(), "frac_dup": frac_dup, "frac_mito": frac_mito},
                       index=pd.Index(bcs))
    genome_len = int(sum(cs.values()))
    frags = DeviceFragments(cell_ptr, pos, val, genome_len, paired=bool(summ.paired))
    key = FRAGMENT_PAIRED if summ.paired else FRAGMENT_SINGLE
    return AnnData(obs=obs, n_obs=R, n_vars=0, obsm={key: frags},
                   uns={"reference_sequences": reference_sequences(cs)})


TEXT_BLOCK_BYTES = 1 << 31  # text per device call when streaming a barcode-sorted file


def _open(path):
    with open(path, "rb") as fh:
        head = fh.read(2)
    return (gzip.open if head == b"\x1f\x8b" else open)(path, "rb")


def _read_text(path) -> bytes:
    with _open(path) as fh:
        return fh.read()


def _barcode(line: bytes):
    """The barcode field of a record line, None for comment / empty / short lines."""
    if not line or line.startswith(b"#"):
        return None
    f = line.rstrip(b"\r").split(b"\t", 4)
    return f[3] if len(f) > 3 else None


def _last_run_start(buf: bytes) -> int:
    """Offset of the first line of the last barcode run among the complete lines
    of buf (0 if that run starts at the beginning, -1 without a complete record)."""
    end = buf.rfind(b"\n")
    if end < 0:
        return -1
    pos, last, start = end, None, -1
    while pos > 0:
        prev = buf.rfind(b"\n", 0, pos)
        line = buf[prev + 1:pos]
        bc = _barcode(line)
        if bc is not None:
            if last is None:
                last = bc
            elif bc != last:
                return start
            start = prev + 1
        pos = prev if prev >= 0 else 0
        if prev < 0:
            break
    return 0 if last is not None else -1


def _text_blocks(path, block_bytes: int):
    """Blocks of a barcode-sorted fragment file, each ending where a barcode run
    ends, so no barcode group is split between two device calls."""
    carry = b""
    with _open(path) as fh:
        while True:
            data = fh.read(block_bytes)
            buf = carry + data
            if not data:
                if buf:
                    yield buf
                return
            cut = _last_run_start(buf)
            if cut <= 0:  # one barcode run fills the buffer: read on
                carry = buf
                continue
            yield buf[:cut]
            carry = buf[cut:]


def _concat(parts):
    """Rows of several imports, in order (fragment stores on the device, obs on the host)."""
    import pandas as pd
    torch = _lib.require_gpu()
    parts = [p for p in parts if p.n_obs > 0]
    if not parts:
        log.warning("No barcodes passed the QC filter. No data is imported.")
        return AnnData(n_obs=0, n_vars=0)
    if len(parts) == 1:
        return parts[0]
    keys = {next(iter(p.obsm)) for p in parts}
    if len(keys) != 1:
        raise _lib.PanicException("paired-end and single-end records in one file")
    key = keys.pop()
    frs = [p.obsm[key] for p in parts]
    offs = np.cumsum([0] + [f.n_fragments for f in frs])
    cell_ptr = torch.cat([frs[0].cell_ptr[:1]] + [f.cell_ptr[1:] + int(o) for f, o in zip(frs, offs[:-1])])
    frags = DeviceFragments(cell_ptr, torch.cat([f.pos for f in frs]), torch.cat([f.values for f in frs]),
                            frs[0].n_cols, frs[0].paired)
    obs = pd.concat([p.obs for p in parts])
    if obs.index.has_duplicates:  # import.rs:139-141
        raise _lib.PanicException("Please sort fragment file by barcodes")
    return AnnData(obs=obs, n_obs=len(obs), n_vars=0, obsm={key: frags}, uns=dict(parts[0].uns))


def import_fragments(
    fragment_file,
    chrom_sizes,
    *,
    file=None,
    min_num_fragments: int = 200,
    sorted_by_barcode: bool = True,
    whitelist=None,
    chrM: list[str] = ["chrM", "M"],
    shift_left: int = 0,
    shift_right: int = 0,
    chunk_size: int = 2000,
    tempdir=None,
    backend="hdf5",
    n_jobs: int = 8,
    _block_bytes: int | None = None,
):
    """Import fragment files and compute basic QC metrics
    (snapatac2.pp.import_fragments, _basic.py:156-340).

    Same arguments and result: obsm['fragment_paired'] (or 'fragment_single'),
    obs 'n_fragment', 'frac_dup', 'frac_mito', uns 'reference_sequences'.
    `chunk_size`, `tempdir` and `n_jobs` only shape the reference's CPU
    processing and do not change the result; backed output (`file`) needs an
    HDF5 library, which this build does not have. A barcode-sorted file is
    streamed through the device in blocks of about TEXT_BLOCK_BYTES ending at
    barcode-run boundaries; an unsorted file is imported in one device call
    (at most 2^31 - 2 lines).
    """
    chrom_sizes = chrom_sizes.chrom_sizes if hasattr(chrom_sizes, "chrom_sizes") else chrom_sizes
    if len(chrom_sizes) == 0:
        raise ValueError("chrom_size cannot be empty")
    if whitelist is not None:
        if isinstance(whitelist, (str, Path)):
            with open(whitelist, "r") as fl:
                whitelist = set(line.strip() for line in fl)
        else:
            whitelist = set(whitelist)
    if isinstance(fragment_file, list):
        if file is not None and len(file) != len(fragment_file):
            raise ValueError("The length of 'file' must be the same as the length of 'fragment_file'")
        return [import_fragments(f, chrom_sizes, file=None if file is None else file[i],
                                 min_num_fragments=min_num_fragments,
                                 sorted_by_barcode=sorted_by_barcode, whitelist=whitelist, chrM=chrM,
                                 shift_left=shift_left, shift_right=shift_right)
                for i, f in enumerate(fragment_file)]
    if file is not None:
        raise NotImplementedError("backed AnnData output needs HDF5, which is not available here")
    kw = dict(min_num_fragments=min_num_fragments, sorted_by_barcode=sorted_by_barcode, whitelist=whitelist,
              chrM=chrM, shift_left=shift_left, shift_right=shift_right)
    if not sorted_by_barcode:  # one device call over the whole text
        return import_fragment_text(_read_text(fragment_file), chrom_sizes, **kw)
    # barcode-sorted input streams through the device in blocks of whole barcode runs
    return _concat([import_fragment_text(b, chrom_sizes, **kw)
                    for b in _text_blocks(fragment_file, _block_bytes or TEXT_BLOCK_BYTES)])


def export_fragment_text(frags: DeviceFragments, chrom_sizes: dict[str, int], barcodes,
                         order=None):
    """Fragments -> text lines as export_fragments writes them
    (snapatac2-core/src/export.rs:90-106), on the device. Line k prints
    fragment order[k] (all fragments row by row if order is None). Returns a
    uint8 CUDA tensor (16-byte padded) and the number of text bytes."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    dev = torch.device("cuda")
    cs = sorted_chrom_sizes(chrom_sizes)
    d_chroms = torch.from_numpy(_name_table(list(cs), list(cs.values()))).to(dev)
    enc = [b.encode() if isinstance(b, str) else bytes(b) for b in barcodes]
    if len(enc) != frags.shape[0]:
        raise ValueError("one barcode per row")
    width = max([len(b) for b in enc] + [1])
    nm = np.zeros((len(enc), width), dtype=np.uint8)
    for i, b in enumerate(enc):
        nm[i, :len(b)] = np.frombuffer(b, dtype=np.uint8)
    d_names = torch.from_numpy(nm).to(dev)
    n = frags.n_fragments
    line_end = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    wsb = L.snapgpu_frag_format_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    d_order = None
    if order is not None:
        order = torch.as_tensor(order, dtype=torch.int64, device=dev).contiguous()
        if order.numel() != n:
            raise ValueError("order must list every fragment once")
        d_order = order.data_ptr()
    total = C.c_int64(0)
    args = (frags.cell_ptr.data_ptr(), frags.shape[0], frags.pos.data_ptr(), frags.values.data_ptr(), n,
            1 if frags.paired else 0, d_chroms.data_ptr(), d_names.data_ptr(), width, d_order,
            line_end.data_ptr())
    _lib.check(L.snapgpu_frag_format(*args, None, 0, ws.data_ptr(), wsb, C.byref(total), st))
    text = torch.empty((total.value + 15) // 16 * 16, dtype=torch.uint8, device=dev)
    _lib.check(L.snapgpu_frag_format(*args, text.data_ptr(), text.numel(), ws.data_ptr(), wsb,
                                     C.byref(total), st))
    return text, total.value
