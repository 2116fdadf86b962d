"""Vibronic-model JSON input/output (the part of pibronic/vibronic/vibronic_model_io.py that
the sampling path reads).

Same file format as the reference: keys from VibronicModelKeys, arrays as nested lists,
coupled shapes E (A,A), w (N,), L (N,A,A), Q (N,N,A,A) and diagonal ("sampling") shapes
E (A,), w (N,), L (N,A), Q (N,N,A) (vibronic_model_io.py:42-73). Arrays that are all zero
are omitted on save and read back as zeros (:517-533, :556-596). Model identity is the
SHA-512 of the file text (:293-331).
"""
import copy
import hashlib
import json
import os

import numpy as np
from numpy import float64 as F64

from .vibronic_model_keys import VibronicModelKeys as VMK


def model_shape_dict(A, N):
    """vibronic_model_io.py:42-56"""
    return {VMK.E: (A, A), VMK.w: (N,), VMK.G1: (N, A, A), VMK.G2: (N, N, A, A),
            VMK.G3: (N, N, N, A, A), VMK.G4: (N, N, N, N, A, A)}


def diagonal_model_shape_dict(A, N):
    """vibronic_model_io.py:59-73"""
    return {VMK.E: (A,), VMK.w: (N,), VMK.G1: (N, A), VMK.G2: (N, N, A),
            VMK.G3: (N, N, N, A), VMK.G4: (N, N, N, N, A)}


def _extract_dimensions_from_dictionary(dictionary):
    """vibronic_model_io.py:473-477 -> (A, N)"""
    return int(dictionary[VMK.A]), int(dictionary[VMK.N])


def _read_json(path):
    assert os.path.isfile(path), f"invalid path:\n{path}"
    with open(path, mode="r", encoding="UTF8") as fh:
        d = json.loads(fh.read())
    VMK.change_dictionary_keys_from_strings_to_enum_members(d)
    return d


def extract_dimensions_of_model(FS=None, path=None):
    """(A, N) of a coupled model file (vibronic_model_io.py:489-498)."""
    if FS is None:
        assert path is not None, "no arguments provided"
    else:
        path = FS.path_vib_model
    return _extract_dimensions_from_dictionary(_read_json(path))


def extract_dimensions_of_diagonal_model(FS=None, path=None):
    """(A, N) of a sampling model file (vibronic_model_io.py:501-514)."""
    if FS is None:
        assert path is not None, "no arguments provided"
    else:
        path = FS.path_rho_model
    return _extract_dimensions_from_dictionary(_read_json(path))


def _verify(kwargs, shape_dict):
    assert VMK.N in kwargs, "need the number of modes"
    assert VMK.A in kwargs, "need the number of surfaces"
    for key, value in kwargs.items():
        if key in shape_dict:
            assert kwargs[key].shape == shape_dict[key], f"{key} have incorrect shape"


def verify_model_parameters(kwargs):
    """vibronic_model_io.py:170-184"""
    A, N = _extract_dimensions_from_dictionary(kwargs)
    _verify(kwargs, model_shape_dict(A, N))


def verify_diagonal_model_parameters(kwargs):
    """vibronic_model_io.py:187-201"""
    A, N = _extract_dimensions_from_dictionary(kwargs)
    _verify(kwargs, diagonal_model_shape_dict(A, N))


def _save_to_JSON(path, dictionary):
    """vibronic_model_io.py:517-533: arrays become lists; all-zero arrays are dropped."""
    d = copy.deepcopy(dictionary)
    VMK.change_dictionary_keys_from_enum_members_to_strings(d)
    for key, value in list(d.items()):
        if isinstance(value, (np.ndarray, np.generic)):
            if np.count_nonzero(value) > 0:
                d[key] = value.tolist()
            else:
                del d[key]
    with open(path, mode="w", encoding="UTF8") as fh:
        fh.write(json.dumps(d))


def save_model_to_JSON(path, dictionary):
    verify_model_parameters(dictionary)
    _save_to_JSON(path, dictionary)


def save_diagonal_model_to_JSON(path, dictionary):
    verify_diagonal_model_parameters(dictionary)
    _save_to_JSON(path, dictionary)


def _load_inplace_from_JSON(path, dictionary):
    """vibronic_model_io.py:556-572: overwrite the given arrays; absent keys become zeros."""
    d = _read_json(path)
    for key, value in dictionary.items():
        if isinstance(value, (np.ndarray, np.generic)):
            if key not in d:
                dictionary[key].fill(0.0)
            else:
                dictionary[key][:] = np.array(d[key], dtype=F64)


def _load_from_JSON(path):
    """vibronic_model_io.py:575-596: lists become float64 arrays; E defaults to zeros."""
    d = _read_json(path)
    for key, value in d.items():
        if isinstance(value, list):
            d[key] = np.array(value, dtype=F64)
    if VMK.E not in d:
        A, N = _extract_dimensions_from_dictionary(d)
        d[VMK.E] = np.zeros(model_shape_dict(A, N)[VMK.E], dtype=F64)
    return d


def load_model_from_JSON(path, dictionary=None):
    """vibronic_model_io.py:599-623"""
    if not bool(dictionary):
        d = _load_from_JSON(path)
        verify_model_parameters(d)
        return d
    verify_model_parameters(dictionary)
    _load_inplace_from_JSON(path, dictionary)
    verify_model_parameters(dictionary)


def load_diagonal_model_from_JSON(path, dictionary=None):
    """vibronic_model_io.py:626-650"""
    if not bool(dictionary):
        d = _load_from_JSON(path)
        verify_diagonal_model_parameters(d)
        return d
    verify_diagonal_model_parameters(dictionary)
    _load_inplace_from_JSON(path, dictionary)
    verify_diagonal_model_parameters(dictionary)


def remove_coupling_from_model(path_source, path_destination):
    """Write the diagonal (in surfaces) part of a coupled model as a sampling model.

    vibronic_model_io.py:749-761: every (..., A, A) array is replaced by its diagonal.
    """
    model = load_model_from_JSON(path_source)
    for key, value in model.items():
        if hasattr(value, "shape") and len(value.shape) >= 2:
            nd = len(value.shape)
            model[key] = np.diagonal(model[key], axis1=nd - 2, axis2=nd - 1).copy()
    save_diagonal_model_to_JSON(path_destination, model)


def _hash(string):
    """SHA-512 hex digest of the text (vibronic_model_io.py:293-297)."""
    m = hashlib.sha512()
    m.update(string.encode("UTF-8"))
    return m.hexdigest()


def _hash_file(path):
    assert os.path.isfile(path), f"The path provided is not a valid file! Path:\n{path}"
    with open(path, mode="r", encoding="UTF8") as fh:
        return _hash(fh.read())


def create_model_hash(FS=None, path=None):
    """vibronic_model_io.py:300-314"""
    if FS is None:
        assert path is not None, "no arguments provided"
    else:
        path = FS.path_vib_model
    return _hash_file(path)


def create_diagonal_model_hash(FS=None, path=None):
    """vibronic_model_io.py:317-331"""
    if FS is None:
        assert path is not None, "no arguments provided"
    else:
        path = FS.path_rho_model
    return _hash_file(path)
