"""Pickle stage files (``myio.py:21-51``): fin1 returns the dict or None on any error;
fout1 writes a dict with the highest protocol.  Only for files this pipeline writes."""
import pickle


def fin1(Filename):
    with open(Filename, "rb") as f:
        try:
            return pickle.load(f)
        except Exception:
            return None


def fout1(Filename, key_list, v_list):
    to_save = dict((key_list[i], v_list[i]) for i in range(len(key_list)))
    with open(Filename, "wb") as f:
        pickle.dump(to_save, f, protocol=pickle.HIGHEST_PROTOCOL)


def fout2(Filename, d):
    fout1(Filename, list(d.keys()), list(d.values()))
