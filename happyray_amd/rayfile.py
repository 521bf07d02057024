"""Ray records: 24 bytes per ray, origin xyz then direction xyz as float32,
back to back with no header -- the input file of
SimpleRayTraverser::loadRaysFromFile (src/RT/Integrator/SimpleRayTraverser.h:200-262)
and the dump of AORayExporter (src/RT/Integrator/AORayExporter.h:262-310).
The same layout is a server request payload (hr_request.layout = HR_RAYS_RECORDS)."""
import numpy as np

RECORD_BYTES = 24


def read_rays(path):
    """(n, 6) float32 records; a trailing partial record is ignored (size / 24)."""
    raw = np.fromfile(path, dtype=np.uint8)
    n = raw.size // RECORD_BYTES
    return raw[:n * RECORD_BYTES].view(np.float32).reshape(n, 6).copy()


def write_rays(path, records):
    np.ascontiguousarray(records, np.float32).reshape(-1, 6).tofile(path)


def to_records(org, dirs):
    """SoA (3, n) origins and directions -> (n, 6) records."""
    o = np.asarray(org, np.float32).reshape(3, -1)
    d = np.asarray(dirs, np.float32).reshape(3, -1)
    return np.ascontiguousarray(np.concatenate([o, d]).T)


def from_records(records):
    """(n, 6) records -> (org (3, n), dirs (3, n))."""
    r = np.asarray(records, np.float32).reshape(-1, 6)
    return np.ascontiguousarray(r[:, 0:3].T), np.ascontiguousarray(r[:, 3:6].T)


def dump_name(frame, alpha):
    """AORayExporter's file name: ray_dump_<frame>_t_<alpha with 2 decimals>.rays."""
    return f"ray_dump_{int(frame)}_t_{float(np.float32(alpha)):.2f}.rays"
