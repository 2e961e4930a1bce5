"""Writes h5_messages.json: how libhdf5 encoded the superblock and the object
header messages of the FP64 datasets in the reference's input file
problems/BP6QD_A_4000/config (written by h5py).  The file is read as bytes, by
the published HDF5 format; nothing of the reference is run.  The writer's test
compares its own encoding against these."""
import json
import os
import struct

SRC = "/root/reference/problems/BP6QD_A_4000/config"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "h5_messages.json")


def main():
    d = open(SRC, "rb").read()
    out = {"superblock_fixed": d[8:24].hex(), "datasets": {}}
    i = 0
    while True:
        i = d.find(b"SNOD", i)
        if i < 0:
            break
        n = struct.unpack_from("<H", d, i + 6)[0]
        for s in range(n):
            oh = struct.unpack_from("<Q", d, i + 8 + 40 * s + 8)[0]
            if oh >= len(d) or d[oh] != 1:
                continue
            nmsg = struct.unpack_from("<H", d, oh + 2)[0]
            p, msgs = oh + 16, []
            for _ in range(nmsg):
                t, sz, fl = struct.unpack_from("<HHB", d, p)
                msgs.append([t, fl, d[p + 8:p + 8 + sz].hex()])
                p += 8 + sz
            types = {m[0]: m for m in msgs}
            if 0x03 not in types or not types[0x03][2].startswith("11"):
                continue  # FP64 datasets only
            rank = int(types[0x01][2][2:4], 16)
            key = f"rank{rank}"
            if key not in out["datasets"]:
                out["datasets"][key] = {"dataspace": types[0x01][1:], "datatype": types[0x03][1:],
                                        "fill": types[0x05][1:], "layout_head": [types[0x08][1],
                                                                                 types[0x08][2][:4]]}
        i += 4
    json.dump(out, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main()
