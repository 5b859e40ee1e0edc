"""Layout helpers and the reference's error metrics."""
import numpy as np


def img2vec(img):
    """Img2vec_Mat2Cpp2D.m:2: single(reshape(fliplr(Img'), [], 1)) for a MATLAB-layout image."""
    img = np.asarray(img)
    return np.fliplr(img.T).reshape(-1, order="F").astype(np.float32)


def vec2img(vec, height, width):
    """Vec2img_Cpp2Mat2D.m:2-3: fliplr(reshape(v, height, width))'."""
    v = np.asarray(vec).reshape((height, width), order="F")
    return np.fliplr(v).T


def rrmse(x, ref):
    """My_toolbox/Evalution_parameter/rRMSE.m:2."""
    x, ref = np.asarray(x, np.float64), np.asarray(ref, np.float64)
    return np.sqrt(np.mean((x - ref) ** 2)) / (np.sqrt(np.mean(ref ** 2)) + np.finfo(float).eps)


def local_error(cur, prev):
    """My_toolbox/Evalution_parameter/LocalError.m:2."""
    cur, prev = np.asarray(cur, np.float64), np.asarray(prev, np.float64)
    return np.linalg.norm(cur - prev) / np.linalg.norm(prev)


def rel_l2(x, ref):
    x, ref = np.asarray(x, np.float64).ravel(), np.asarray(ref, np.float64).ravel()
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)
