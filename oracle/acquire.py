"""acquire_all_* result mapping and CombinedOutput JSON (oracle restatement).

Follows acquisition.rs:304-395 (and the five siblings), lib.rs:31-45 and main.rs:242-255:
per SV the per-antenna strengths (best_peak as f64), phases ((best_codephase mod P)/P),
freqs (fc[best] - center), cn0_acr only when some antenna has an estimate, median/MAD fields
only with more than one antenna, results sorted by the sv label, MAD filters dropping SVs whose
MAD exceeds the threshold (a missing MAD passes).
"""
from . import acr as _acr
from . import codes as _codes
from . import jsonfmt as _jsonfmt
from . import stats as _stats

KEY_ORDER = ["gps", "galileo", "beidou", "sbas", "l1c", "qzss"]  # lib.rs:33-44


def system_output(system, prns, peaks, fc, center, P, antenna_numbers, cn0,
                  filter_phase_mad=None, filter_freq_mad=None):
    """peaks[a][j]: dict with best_peak, second_peak, best_freq_idx, best_codephase."""
    label = _codes.SYSTEMS[system][2]
    table = _acr.TABLES[_codes.SYSTEMS[system][3]]
    results = []
    n_ant = len(antenna_numbers)
    for j, prn in enumerate(prns):
        strengths, phases, freqs, seconds = [], [], [], []
        for a in range(n_ant):
            pk = peaks[a][j]
            strengths.append(float(pk["best_peak"]))
            seconds.append(float(pk["second_peak"]))
            phases.append((int(pk["best_codephase"]) % P) / float(P))
            freqs.append(fc[int(pk["best_freq_idx"])] - center)
        rec = {"sv": label.format(prn), "strengths": strengths, "phases": phases,
               "freqs": freqs}
        if cn0:
            cn = [_acr.acr_from_peaks(vm, vs, table) for vm, vs in zip(strengths, seconds)]
            if any(c is not None for c in cn):
                rec["cn0_acr"] = cn
        if n_ant > 1:
            pm = _stats.median(phases)
            fm = _stats.median(freqs)
            rec["phase_median"] = pm
            rec["phase_mad"] = _stats.mad(phases, pm)
            rec["freq_median"] = fm
            rec["freq_mad"] = _stats.mad(freqs, fm)
        if filter_phase_mad is not None and "phase_mad" in rec and not (
                rec["phase_mad"] <= filter_phase_mad):
            continue
        if filter_freq_mad is not None and "freq_mad" in rec and not (
                rec["freq_mad"] <= filter_freq_mad):
            continue
        results.append(rec)
    results.sort(key=lambda r: r["sv"].encode())
    return {"antenna_numbers": list(antenna_numbers), "results": results}


def combined_json(outputs):
    """outputs: {system: system_output(...)} -> serde_json pretty text of CombinedOutput."""
    ordered = {k: outputs[k] for k in KEY_ORDER if k in outputs}
    return _jsonfmt.dumps_pretty(ordered)
