"""Reduce rocprofv3 CSV output (``--output-format csv``) to short text summaries.

    python -m gmaster_amd.profsum stats <kernel_stats.csv>
        per-kernel calls / total / average duration
    python -m gmaster_amd.profsum pmc <counter_collection.csv> COUNTER
        per-kernel dispatch count, mean and max of one counter
    python -m gmaster_amd.profsum traffic <fetch.csv> <write.csv> KERNEL COPY_BYTES
        HBM bytes per launch of KERNEL from FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
        each scaled by what the same counter reads for the largest device copy kernel
        dispatch of COPY_BYTES read + COPY_BYTES written in the same run
    python -m gmaster_amd.profsum traffic-summaries <fetch.txt> <write.txt> KERNEL COPY_BYTES
        the same from two ``pmc`` summaries
"""
import csv
import re
import sys
from collections import defaultdict

_KERNELS = ("k_coupling_pair_values", "k_coupling_rows", "k_toeplitz_correct", "k_bat_ana_h", "k_bat_ana", "k_bat_syn64", "k_bat_syn", "k_bat_finish", "k_bat_prep", "k_ana_s0", "k_ana_s", "k_syn_s0_mx", "k_syn_s_mx", "k_syn_s0", "k_syn_s", "k_coupling_pairs",
            "k_coupling", "k_toeplitz", "k_ring_fwd_s0", "k_ring_fwd_s", "k_ring_inv_s0", "k_ring_inv_s", "k_fold_s0", "k_fold_s",
            "k_finish_s0", "k_finish_s", "k_prep_s0", "k_prep_s", "k_lat_tables", "k_cl_partial", "k_cl_final",
            "k_bin_block", "k_bin_cell", "k_gauss_legendre", "k_wigner_d_table", "k_wigner_d_corr",
            "k_table_combine", "k_parity_split", "k_ring_table", "k_masked_links", "k_apodize", "k_gaussian_alms",
            "k_spin_mul", "k_alm_filter", "k_dot_partial", "k_dot_final", "k_deproject", "k_lu_clear", "k_lu_diag",
            "k_lu_panel", "k_lu_update", "k_lu_solve", "k_coupling_pair_values", "k_toeplitz_correct", "k_fingerprint")


def short(name):
    """``k_ana_s<8,4>``-style name of one of this package's kernels (mangled or demangled),
    else the first 60 characters."""
    for key in _KERNELS:
        m = re.search(r"(?<![A-Za-z_])" + key + r"(<[^>]*>|\(|I|E|$)", name)
        if not m:
            continue
        tail = m.group(1)
        if tail.startswith("<"):
            return key + tail.replace(" ", "")
        if tail == "I":
            tmpl = name[m.end() - 1:]
            return key + "<" + tmpl[1:].split("EEEv")[0].replace("ILi", "").replace("ELi", ",").replace("Li", "") + ">"
        return key
    return name[:60]


def stats(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print(f"{'kernel':40s} {'calls':>6s} {'total_ms':>10s} {'avg_ms':>10s} {'pct':>6s}")
    for r in rows:
        print(f"{short(r['Name']):40s} {int(r['Calls']):6d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
              f"{float(r['AverageNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")


def _per_dispatch(path, counter):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name", counter) != counter:
            continue
        acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return acc


def _first_copy(path, counter):
    """Counter value of the first runtime copy dispatch of at least 64 MiB in dispatch order:
    bench.py's ``--calibrate-copy`` 1 GiB copy, which runs before any pipeline work (small
    table uploads come earlier; later copies, such as clones of map inputs, can be larger)."""
    best = None
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name", counter) != counter or "copybuffer" not in r["Kernel_Name"].lower():
            continue
        if float(r["Counter_Value"]) < 64 * 1024:
            continue
        d = int(r.get("Dispatch_Id", 0))
        if best is None or d < best[0]:
            best = (d, float(r["Counter_Value"]))
    if best is None:
        raise SystemExit("no copy kernel dispatch in the counter file")
    return best[1]


def pmc(path, counter):
    acc = defaultdict(list)
    for k, v in _per_dispatch(path, counter).items():
        acc[short(k)] += v
    print(f"{'kernel':40s} {'dispatches':>10s} {'mean_KiB':>16s} {'max_KiB':>16s}   ({counter})")
    for k, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k:40s} {len(v):10d} {sum(v) / len(v):16.1f} {max(v):16.1f}")
    try:
        print(f"# calibration copy (first copy dispatch): {_first_copy(path, counter):.1f} KiB")
    except SystemExit:
        pass


def _is(name, kernel):
    """``kernel`` is a bare name (any template arguments) or ``name<args>`` (exact)."""
    return name == kernel if "<" in kernel else name.split("<")[0] == kernel


def traffic(fetch_path, write_path, kernel, copy_bytes):
    copy_bytes = float(copy_bytes)
    out = {}
    for tag, path, counter in (("fetch", fetch_path, "FETCH_SIZE"), ("write", write_path, "WRITE_SIZE")):
        acc = _per_dispatch(path, counter)
        ref_kib = _first_copy(path, counter)
        scale = copy_bytes / (ref_kib * 1024.0)
        vals = [x for k, v in acc.items() if _is(short(k), kernel) for x in v]
        if not vals:
            raise SystemExit(f"kernel {kernel} not found in {path}")
        mean_kib = sum(vals) / len(vals)
        out[tag] = mean_kib * 1024.0 * scale
        print(f"{counter}: {kernel} mean {mean_kib:.1f} KiB over {len(vals)} dispatches; copy kernel "
              f"{ref_kib:.1f} KiB for {copy_bytes / 2**30:.3f} GiB -> scale {scale:.4f}; "
              f"scaled {out[tag] / 1e9:.3f} GB per launch")
    print(f"traffic_bytes_per_launch: {out['fetch'] + out['write']:.6e}")


def _summary_rows(path):
    rows = []
    for line in open(path):
        if line.startswith("#") or line.startswith("kernel") or not line.strip():
            continue
        parts = line.rstrip("\n").rsplit(None, 3)
        rows.append((parts[0].strip(), int(parts[1]), float(parts[2]), float(parts[3])))
    return rows


def traffic_from_summaries(fetch_txt, write_txt, kernel, copy_bytes):
    """As ``traffic``, from two ``pmc`` summaries (mean per kernel; the calibration copy line)."""
    copy_bytes = float(copy_bytes)
    out = {}
    for tag, path, counter in (("fetch", fetch_txt, "FETCH_SIZE"), ("write", write_txt, "WRITE_SIZE")):
        rows = _summary_rows(path)
        cal = [float(line.split(":", 1)[1].split()[0]) for line in open(path) if line.startswith("# calibration copy")]
        if not cal:
            raise SystemExit(f"{path} has no calibration copy line")
        ref_kib = cal[0]
        hits = [r for r in rows if _is(short(r[0]), kernel)]
        if not hits:
            raise SystemExit(f"kernel {kernel} not found in {path}")
        # a bare name covers every instantiation: dispatch-weighted mean
        n = sum(h[1] for h in hits)
        mean_kib = sum(h[1] * h[2] for h in hits) / n
        name = kernel if len(hits) > 1 else short(hits[0][0])
        scale = copy_bytes / (ref_kib * 1024.0)
        out[tag] = mean_kib * 1024.0 * scale
        print(f"{counter}: {name} mean {mean_kib:.1f} KiB over {n} dispatches; copy kernel "
              f"{ref_kib:.1f} KiB for {copy_bytes / 2**30:.3f} GiB -> scale {scale:.4f}; "
              f"scaled {out[tag] / 1e9:.3f} GB per launch")
    print(f"traffic_bytes_per_launch: {out['fetch'] + out['write']:.6e}")


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "stats":
        stats(sys.argv[2])
    elif cmd == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    elif cmd == "traffic":
        traffic(*sys.argv[2:6])
    elif cmd == "traffic-summaries":
        traffic_from_summaries(*sys.argv[2:6])
    else:
        raise SystemExit(__doc__)
