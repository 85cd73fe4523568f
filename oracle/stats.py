"""Robust statistics (oracle restatement of stats.rs:10-30)."""


def median(data):
    if len(data) == 0:
        raise ValueError("median of empty slice")
    c = sorted(data)
    mid = len(c) // 2
    return (c[mid - 1] + c[mid]) / 2.0 if len(c) % 2 == 0 else c[mid]


def mad(data, med):
    if len(data) == 0:
        raise ValueError("MAD of empty slice")
    return median([abs(x - med) for x in data])
