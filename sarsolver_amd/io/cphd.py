"""Compensated Phase History Data (CPHD) files, versions 1.0.x and 0.3, read
without sarpy.

The reference reads CPHD through sarpy's phase-history readers
(io/cphd101.py:13-285, io/cphd03.py:13-289: `open_phase_history`,
`read_pvp_variable`, `read`).  sarpy is not available here, so this module
parses the published layout directly:

* CPHD 1.0.x (NGA.STND.0068-1): a text header "CPHD/1.0.x" followed by
  "KEY := value" lines and a form feed; an XML block; a per-vector parameter
  (PVP) block whose fields are described in the XML <PVP> element by
  <Offset> and <Size> in 8-byte words and a <Format> such as "F8" or
  "X=F8;Y=F8;Z=F8;"; a signal block of complex samples ("CF8": two
  big-endian float32, "CI4": two int16, "CI2": two int8) per channel at
  <SignalArrayByteOffset>, <PVPArrayByteOffset> within the blocks.
* CPHD 0.3: header "CPHD/0.3" with XML_BYTE_OFFSET / VB_BYTE_OFFSET /
  CPHD_BYTE_OFFSET; the per-vector block (VBP) holds the fields of the XML
  <VectorParameters> element in document order, each element's text being
  its size in bytes; sample types "RE32F_IM32F", "RE16I_IM16I",
  "RE08I_IM08I"; channels stored one after another.

Binary data are big-endian.  When a per-vector amplitude scale factor
(AmpSF) is present, signal values are scaled by it.  Compressed signal
arrays are not supported.
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

# Signal sample formats: (numpy component dtype, bytes per complex sample, library code)
SAMPLE_FORMATS = {
    "CF8": (">f4", 8, 0), "CI4": (">i2", 4, 1), "CI2": (">i1", 2, 2),
    "RE32F_IM32F": (">f4", 8, 0), "RE16I_IM16I": (">i2", 4, 1), "RE08I_IM08I": (">i1", 2, 2),
}

_PVP_SCALARS = {"F8": ">f8", "I8": ">i8", "U8": ">u8", "F4": ">f4", "I4": ">i4", "U4": ">u4"}


def _strip_ns(root: ET.Element) -> ET.Element:
    for el in root.iter():
        if isinstance(el.tag, str) and "}" in el.tag:
            el.tag = el.tag.split("}", 1)[1]
    return root


def _text(el: Optional[ET.Element], path: str, cast=str, default=None):
    node = el.find(path) if el is not None else None
    if node is None or node.text is None:
        if default is not None:
            return default
        raise ValueError(f"CPHD XML: missing <{path}>")
    return cast(node.text.strip())


def read_header(f) -> (str, Dict[str, str], int):
    """(version, key/value pairs, byte length of the header) of a CPHD file."""
    first = f.readline()
    m = re.match(rb"CPHD/(\d+\.\d+(?:\.\d+)?)\s*$", first)
    if not m:
        raise ValueError(f"not a CPHD file (first line {first[:32]!r})")
    version = m.group(1).decode()
    kv = {}
    while True:
        line = f.readline()
        if not line:
            raise ValueError("CPHD header: no section terminator")
        if line in (b"\f\n", b"\x0c\n"):
            break
        k, sep, v = line.decode("ascii", errors="replace").partition(":=")
        if not sep:
            raise ValueError(f"CPHD header: bad line {line[:64]!r}")
        kv[k.strip()] = v.strip()
    return version, kv, f.tell()


@dataclass
class ChannelLayout:
    identifier: str
    num_vectors: int
    num_samples: int
    signal_offset: int      # absolute byte offset of the channel's signal array
    pvp_offset: int         # absolute byte offset of the channel's per-vector parameters


@dataclass
class CphdFile:
    """Parsed layout of a CPHD file; arrays are read on demand."""
    path: Path
    version: str
    header: Dict[str, str]
    xml: ET.Element
    sample_format: str
    bytes_per_vector_params: int
    pvp_dtype: np.dtype
    channels: List[ChannelLayout] = field(default_factory=list)
    sign: int = -1

    @property
    def component_dtype(self) -> str:
        return SAMPLE_FORMATS[self.sample_format][0]

    @property
    def sample_bytes(self) -> int:
        return SAMPLE_FORMATS[self.sample_format][1]

    @property
    def format_code(self) -> int:
        return SAMPLE_FORMATS[self.sample_format][2]

    def pvp(self, channel: int = 0) -> np.ndarray:
        """Structured array of the channel's per-vector parameters."""
        ch = self.channels[channel]
        return np.fromfile(self.path, dtype=self.pvp_dtype, count=ch.num_vectors,
                           offset=ch.pvp_offset)

    def pvp_field(self, name: str, channel: int = 0, index=slice(None)) -> np.ndarray:
        """One parameter as native-endian FP64 (vectors [N] or [N, 3])."""
        p = self.pvp(channel)
        if name not in p.dtype.names:
            raise KeyError(f"CPHD file has no per-vector parameter {name!r}")
        v = p[name]
        if v.dtype.names:  # X;Y;Z composite
            v = np.stack([v[n] for n in v.dtype.names], axis=-1)
        return np.asarray(v, dtype=np.float64)[index]

    def raw_signal(self, channel: int = 0, start: int = 0, stop: Optional[int] = None) -> np.ndarray:
        """Raw big-endian bytes of vectors [start, stop) of a channel: uint8
        array [n, num_samples * sample_bytes]."""
        ch = self.channels[channel]
        stop = ch.num_vectors if stop is None else min(stop, ch.num_vectors)
        n = max(0, stop - start)
        row = ch.num_samples * self.sample_bytes
        # copy-on-write mapping: a writable view without reading the file twice
        mm = np.memmap(self.path, dtype=np.uint8, mode="c", offset=ch.signal_offset + start * row,
                       shape=(n * row,)) if n else np.zeros(0, np.uint8)
        return np.asarray(mm).reshape(n, row)

    def amp_sf(self, channel: int = 0, start: int = 0, stop: Optional[int] = None):
        p = self.pvp(channel)
        if "AmpSF" not in p.dtype.names:
            return None
        return np.asarray(p["AmpSF"], dtype=np.float64)[start:stop]

    def signal(self, channel: int = 0, index=slice(None)) -> np.ndarray:
        """Decoded complex64 samples [n, num_samples] of the selected vectors
        (host decode; AmpSF applied when present)."""
        ch = self.channels[channel]
        start, stop, step = index.indices(ch.num_vectors) if isinstance(index, slice) else (index, index + 1, 1)
        raw = self.raw_signal(channel, start, stop)[::step]
        comp = raw.view(self.component_dtype).reshape(raw.shape[0], ch.num_samples, 2)
        re = comp[..., 0].astype(np.float32)
        im = comp[..., 1].astype(np.float32)
        sf = self.amp_sf(channel, start, stop)
        if sf is not None:  # each component times the vector's scale (FP32)
            s32 = sf[::step, None].astype(np.float32)
            re, im = re * s32, im * s32
        out = np.empty(re.shape, dtype=np.complex64)
        out.real, out.imag = re, im
        return out


def _pvp_dtype_101(pvp: ET.Element, nbytes: int) -> np.dtype:
    names, formats, offsets = [], [], []

    def add(tag, el):
        off = _text(el, "Offset", int) * 8
        size = _text(el, "Size", int)
        fmt = _text(el, "Format", str)
        if fmt in _PVP_SCALARS:
            dt = np.dtype(_PVP_SCALARS[fmt])
        else:
            parts = [p for p in fmt.split(";") if p]
            sub = [(p.split("=")[0], _PVP_SCALARS[p.split("=")[1]]) for p in parts]
            dt = np.dtype(sub)
        if dt.itemsize > size * 8:
            raise ValueError(f"PVP {tag}: format {fmt} exceeds {size} words")
        names.append(tag)
        formats.append(dt)
        offsets.append(off)

    for el in pvp:
        if el.tag == "AddedPVP":
            add(_text(el, "Name"), el)
        elif el.find("Offset") is not None:
            add(el.tag, el)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": nbytes})


def _open_101(path: Path, version: str, kv: Dict[str, str], xml: ET.Element) -> CphdFile:
    data = xml.find("Data")
    fmt = _text(data, "SignalArrayFormat")
    if fmt not in ("CF8", "CI4", "CI2"):
        raise ValueError(f"unsupported SignalArrayFormat {fmt}")
    nbytes = _text(data, "NumBytesPVP", int)
    pvp_base = int(kv["PVP_BLOCK_BYTE_OFFSET"])
    sig_base = int(kv["SIGNAL_BLOCK_BYTE_OFFSET"])
    chans = []
    for ch in data.findall("Channel"):
        if ch.find("CompressedSignalSize") is not None:
            raise ValueError("compressed CPHD signal arrays are not supported")
        chans.append(ChannelLayout(_text(ch, "Identifier"), _text(ch, "NumVectors", int),
                                   _text(ch, "NumSamples", int),
                                   sig_base + _text(ch, "SignalArrayByteOffset", int),
                                   pvp_base + _text(ch, "PVPArrayByteOffset", int)))
    sgn = _text(xml, "Global/SGN", int)
    return CphdFile(path, version, kv, xml, fmt, nbytes, _pvp_dtype_101(xml.find("PVP"), nbytes),
                    chans, sgn)


_VBP_XYZ = {"TxPos", "RcvPos", "SRPPos", "TxVel", "RcvVel"}


def _open_03(path: Path, version: str, kv: Dict[str, str], xml: ET.Element) -> CphdFile:
    data = xml.find("Data")
    fmt = _text(data, "SampleType")
    if fmt not in ("RE32F_IM32F", "RE16I_IM16I", "RE08I_IM08I"):
        raise ValueError(f"unsupported SampleType {fmt}")
    nbytes = _text(data, "NumBytesVBP", int)
    names, formats, offsets = [], [], []
    off = 0
    for el in xml.find("VectorParameters").iter():
        if len(el) or el.text is None or not el.text.strip().isdigit():
            continue
        size = int(el.text.strip())
        if el.tag in _VBP_XYZ:
            dt = np.dtype([("X", ">f8"), ("Y", ">f8"), ("Z", ">f8")])
        else:
            dt = np.dtype(">f8")
        if dt.itemsize != size:
            raise ValueError(f"VBP {el.tag}: size {size} does not match its type")
        names.append(el.tag)
        formats.append(dt)
        offsets.append(off)
        off += size
    if off > nbytes:
        raise ValueError("VBP fields exceed NumBytesVBP")
    vdt = np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": nbytes})
    vb_base = int(kv["VB_BYTE_OFFSET"])
    sig_base = int(kv["CPHD_BYTE_OFFSET"])
    chans = []
    vb_off, sig_off = vb_base, sig_base
    sbytes = SAMPLE_FORMATS[fmt][1]
    for i, a in enumerate(data.findall("ArraySize")):
        nv, ns = _text(a, "NumVectors", int), _text(a, "NumSamples", int)
        chans.append(ChannelLayout(str(i), nv, ns, sig_off, vb_off))
        vb_off += nv * nbytes
        sig_off += nv * ns * sbytes
    sgn = _text(xml, "Global/PhaseSGN", int)
    return CphdFile(path, version, kv, xml, fmt, nbytes, vdt, chans, sgn)


def open_cphd(path) -> CphdFile:
    """Parse the header and XML of a CPHD 1.0.x or 0.3 file."""
    path = Path(path)
    with open(path, "rb") as f:
        version, kv, _ = read_header(f)
        if version.startswith("1."):
            xoff, xsize = int(kv["XML_BLOCK_BYTE_OFFSET"]), int(kv["XML_BLOCK_SIZE"])
        elif version.startswith("0.3"):
            xoff, xsize = int(kv["XML_BYTE_OFFSET"]), int(kv["XML_DATA_SIZE"])
        else:
            raise ValueError(f"unsupported CPHD version {version}")
        f.seek(xoff)
        try:
            xml = _strip_ns(ET.fromstring(f.read(xsize)))
        except ET.ParseError as e:
            raise ValueError(f"CPHD XML block does not parse: {e}") from e
    if version.startswith("1."):
        return _open_101(path, version, kv, xml)
    return _open_03(path, version, kv, xml)
