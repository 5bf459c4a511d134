import ctypes as C, sys, os, time
sys.path.insert(0, os.getcwd())
import av1gpu_amd as A
from av1gpu_amd import learn
from concurrent.futures import ThreadPoolExecutor
L = A.lib()
L.av1mi_prof_read.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
spec = A.batch_spec(256)
with ThreadPoolExecutor(16) as ex:
    bplanes = list(ex.map(lambda s: A.synth_image(s[3], s[0], s[1]), spec))
frames = [A.DeviceFrame(p, W, H, qindex=q) for p, (W, H, q, _) in zip(bplanes, spec)]
enc = A.GpEncoder().init(A.GpConfig(qindex=120))
learn.set_shortlist(enc, learn.load_params(), 0.8)
buf = (C.c_ulonglong * 16)()
for rep in range(2):
    L.av1mi_prof_read(buf, 1)
    out, t = enc.encode(frames, want_recon=False)
    print("recon ms", t.recon, flush=True)
L.av1mi_prof_read(buf, 0)
names = ["luma setup", "luma residuals", "luma tasks", "luma final", "chroma setup", "chroma CfL", "chroma screening+topK",
         "chroma tasks", "chroma final", "leaf maps/records"]
tot = sum(buf[k] for k in range(10))
for k in range(10):
    print("%-24s %6.1f%%" % (names[k], 100.0 * buf[k] / tot))
