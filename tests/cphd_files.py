"""Writers of small CPHD 1.0.1 and 0.3 files for the reader tests (test
infrastructure).  The layouts are those sarsolver_amd/io/cphd.py documents:
text header with byte offsets, namespaced XML, big-endian per-vector
parameters and signal samples."""
from __future__ import annotations

from pathlib import Path

import numpy as np

_COMP = {"CF8": ">f4", "CI4": ">i2", "CI2": ">i1",
         "RE32F_IM32F": ">f4", "RE16I_IM16I": ">i2", "RE08I_IM08I": ">i1"}


def _encode(signal: np.ndarray, fmt: str) -> bytes:
    dt = _COMP[fmt]
    P, N = signal.shape
    comp = np.empty((P, N, 2), dtype=dt)
    if dt == ">f4":
        comp[..., 0], comp[..., 1] = signal.real, signal.imag
    else:
        re, im = np.rint(signal.real), np.rint(signal.imag)
        assert np.array_equal(re, signal.real) and np.array_equal(im, signal.imag)
        comp[..., 0], comp[..., 1] = re, im
    return comp.tobytes()


def _pad(buf: bytearray, align: int = 64) -> int:
    while len(buf) % align:
        buf.append(0)
    return len(buf)


def _assemble(first_line: str, keys, xml: bytes, blocks) -> bytes:
    """Header lines with offsets resolved (until their digit counts settle),
    then the XML and the binary blocks, each 64-byte aligned."""
    offsets = {k: 0 for k in keys}
    for _ in range(4):
        head = first_line + "\n" + "".join(f"{k} := {offsets[k]}\n" for k in keys) + \
            "CLASSIFICATION := UNCLASSIFIED\nRELEASE_INFO := UNRESTRICTED\n\f\n"
        buf = bytearray(head.encode("ascii"))
        new = dict(offsets)
        xoff = _pad(buf)
        buf += xml + b"\f\n"
        new[keys[0]], new[keys[1]] = len(xml), xoff
        for i, blk in enumerate(blocks):
            off = _pad(buf)
            buf += blk
            new[keys[2 + 2 * i]], new[keys[3 + 2 * i]] = len(blk), off
        if new == offsets:
            return bytes(buf)
        offsets = new
    raise RuntimeError("header offsets did not settle")


def write_cphd101(path, channels, sgn: int = -1, fmt: str = "CF8", compressed: bool = False) -> Path:
    """channels: list of dicts with signal [P, N], tx, rx, srp [P, 3], sc0, scss
    (scalars or [P]) and optional amp_sf [P].  `compressed` declares a
    CompressedSignalSize (the samples stay uncompressed)."""
    use_sf = any(ch.get("amp_sf") is not None for ch in channels)
    fields = [("TxTime", 1, "F8"), ("TxPos", 3, "X=F8;Y=F8;Z=F8;"), ("RcvTime", 1, "F8"),
              ("RcvPos", 3, "X=F8;Y=F8;Z=F8;"), ("SRPPos", 3, "X=F8;Y=F8;Z=F8;"),
              ("SC0", 1, "F8"), ("SCSS", 1, "F8")] + ([("AmpSF", 1, "F8")] if use_sf else [])
    nwords = sum(f[1] for f in fields) + 1  # one spare word
    pvp_xml, off = [], 0
    for name, size, f in fields:
        pvp_xml.append(f"<{name}><Offset>{off}</Offset><Size>{size}</Size><Format>{f}</Format></{name}>")
        off += size
    pvp_blk, sig_blk, chan_xml = bytearray(), bytearray(), []
    for i, ch in enumerate(channels):
        sig = np.asarray(ch["signal"])
        P, N = sig.shape
        rec = np.zeros((P, nwords), dtype=">f8")
        col = 0
        vals = {"TxTime": np.arange(P) * 1e-3, "TxPos": ch["tx"], "RcvTime": np.arange(P) * 1e-3,
                "RcvPos": ch["rx"], "SRPPos": ch["srp"],
                "SC0": np.broadcast_to(ch["sc0"], (P,)), "SCSS": np.broadcast_to(ch["scss"], (P,)),
                "AmpSF": np.ones(P) if ch.get("amp_sf") is None else ch["amp_sf"]}
        for name, size, _ in fields:
            rec[:, col:col + size] = np.asarray(vals[name], dtype=np.float64).reshape(P, size)
            col += size
        chan_xml.append(f"<Channel><Identifier>ch{i}</Identifier><NumVectors>{P}</NumVectors>"
                        f"<NumSamples>{N}</NumSamples><SignalArrayByteOffset>{len(sig_blk)}"
                        f"</SignalArrayByteOffset><PVPArrayByteOffset>{len(pvp_blk)}"
                        f"</PVPArrayByteOffset>"
                        + (f"<CompressedSignalSize>{sig.size}</CompressedSignalSize>" if compressed else "")
                        + "</Channel>")
        pvp_blk += rec.tobytes()
        sig_blk += _encode(sig, fmt)
    xml = ('<?xml version="1.0"?><CPHD xmlns="http://api.nsgreg.nga.mil/schema/cphd/1.0.1">'
           f"<Global><DomainType>FX</DomainType><SGN>{sgn:+d}</SGN></Global>"
           f"<Data><SignalArrayFormat>{fmt}</SignalArrayFormat><NumBytesPVP>{nwords * 8}</NumBytesPVP>"
           f"<NumCPHDChannels>{len(channels)}</NumCPHDChannels>{''.join(chan_xml)}"
           "<NumSupportArrays>0</NumSupportArrays></Data>"
           f"<PVP>{''.join(pvp_xml)}</PVP></CPHD>").encode()
    keys = ["XML_BLOCK_SIZE", "XML_BLOCK_BYTE_OFFSET", "PVP_BLOCK_SIZE", "PVP_BLOCK_BYTE_OFFSET",
            "SIGNAL_BLOCK_SIZE", "SIGNAL_BLOCK_BYTE_OFFSET"]
    data = _assemble("CPHD/1.0.1", keys, xml, [bytes(pvp_blk), bytes(sig_blk)])
    path = Path(path)
    path.write_bytes(data)
    return path


def write_cphd03(path, channels, sgn: int = -1, fmt: str = "RE32F_IM32F") -> Path:
    """As write_cphd101 with fx0 / fx_ss in place of sc0 / scss."""
    use_sf = any(ch.get("amp_sf") is not None for ch in channels)
    fields = [("TxTime", 1), ("TxPos", 3), ("RcvTime", 1), ("RcvPos", 3), ("SRPTime", 1),
              ("SRPPos", 3)] + ([("AmpSF", 1)] if use_sf else []) + \
        [("Fx0", 1), ("Fx_SS", 1), ("Fx1", 1), ("Fx2", 1)]
    nwords = sum(f[1] for f in fields)
    vb_blk, sig_blk, sizes = bytearray(), bytearray(), []
    for ch in channels:
        sig = np.asarray(ch["signal"])
        P, N = sig.shape
        rec = np.zeros((P, nwords), dtype=">f8")
        fx0 = np.broadcast_to(ch["fx0"], (P,))
        fss = np.broadcast_to(ch["fx_ss"], (P,))
        vals = {"TxTime": np.arange(P) * 1e-3, "TxPos": ch["tx"], "RcvTime": np.arange(P) * 1e-3,
                "RcvPos": ch["rx"], "SRPTime": np.arange(P) * 1e-3, "SRPPos": ch["srp"],
                "AmpSF": np.ones(P) if ch.get("amp_sf") is None else ch["amp_sf"],
                "Fx0": fx0, "Fx_SS": fss, "Fx1": fx0, "Fx2": fx0 + N * fss}
        col = 0
        for name, size in fields:
            rec[:, col:col + size] = np.asarray(vals[name], dtype=np.float64).reshape(P, size)
            col += size
        vb_blk += rec.tobytes()
        sig_blk += _encode(sig, fmt)
        sizes.append(f"<ArraySize><NumVectors>{P}</NumVectors><NumSamples>{N}</NumSamples></ArraySize>")
    vp = []
    for name, size in fields:
        if name in ("Fx0", "Fx_SS", "Fx1", "Fx2"):
            continue
        vp.append(f"<{name}>{size * 8}</{name}>")
    vp.append("<FxParameters><Fx0>8</Fx0><Fx_SS>8</Fx_SS><Fx1>8</Fx1><Fx2>8</Fx2></FxParameters>")
    xml = ('<?xml version="1.0"?><CPHD xmlns="urn:CPHD:0.3">'
           f"<Data><SampleType>{fmt}</SampleType><NumCPHDChannels>{len(channels)}</NumCPHDChannels>"
           f"<NumBytesVBP>{nwords * 8}</NumBytesVBP>{''.join(sizes)}</Data>"
           f"<Global><DomainType>FX</DomainType><PhaseSGN>{sgn:+d}</PhaseSGN></Global>"
           f"<VectorParameters>{''.join(vp)}</VectorParameters></CPHD>").encode()
    keys = ["XML_DATA_SIZE", "XML_BYTE_OFFSET", "VB_DATA_SIZE", "VB_BYTE_OFFSET",
            "CPHD_DATA_SIZE", "CPHD_BYTE_OFFSET"]
    data = _assemble("CPHD/0.3", keys, xml, [bytes(vb_blk), bytes(sig_blk)])
    path = Path(path)
    path.write_bytes(data)
    return path
