"""Helpers shared by the GPU parity tests."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[gpu-test]", *a, flush=True)


def load_liblz4_fixture():
    z = np.load(os.path.join(HERE, "golden", "liblz4_pages.npz"))
    sizes = z["sizes"]
    blob = z["blob"]
    offs = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))]).astype(np.int64)
    pages = [blob[offs[i]:offs[i + 1]].tobytes() for i in range(sizes.size)]
    return z["data"], int(z["page_rows"]), pages


def q1_dataset(orc, n, seed=11, n_days=2556):
    """Config-1 distributions: date_key over n_days consecutive days from 1992-01-01, discount
    U[0,10], quantity U[1,50], price_cents U[90000,10500000]; date dim (key, year, month)."""
    key, year, month = orc.gen_date_dim(n_days)
    date, disc, qty, price = orc.gen_q1_fact(seed, n, key)
    return dict(d_key=key, d_year=year, d_month=month, date=date, disc=disc, qty=qty, price=price)
