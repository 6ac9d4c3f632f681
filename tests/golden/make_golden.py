"""Writes tests/golden/*.json: known-answer data transcribed from the
reference's unit tests (hmat_lib tests/unit_tests/*.c).  Each entry cites
the file:line it comes from.  Data only (inputs and expected outputs)."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def block_sizes():
    # test_constructors.c:149-254 (halves; first entry = m) and 342-436
    # (custom leaf sizes), 451-470 (uneven custom trees)
    halves = [
        [1, [8, 4, 4]], [1, [9, 5, 4]], [1, [11, 6, 5]], [1, [13, 7, 6]],
        [2, [8, 4, 4, 2, 2, 2, 2]], [2, [9, 5, 4, 3, 2, 2, 2]],
        [2, [11, 6, 5, 3, 3, 3, 2]], [2, [13, 7, 6, 4, 3, 3, 3]],
        [1, [139, 70, 69]], [1, [597, 299, 298]], [1, [2048, 1024, 1024]],
        [2, [139, 70, 69, 35, 35, 35, 34]],
        [2, [597, 299, 298, 150, 149, 149, 149]],
        [3, [139, 70, 69, 35, 35, 35, 34, 18, 17, 18, 17, 18, 17, 17, 17]],
        [3, [597, 299, 298, 150, 149, 149, 149, 75, 75, 75, 74, 75, 74, 75, 74]],
        [4, [139, 70, 69, 35, 35, 35, 34, 18, 17, 18, 17, 18, 17, 17, 17,
             9, 9, 9, 8, 9, 9, 9, 8, 9, 9, 9, 8, 9, 8, 9, 8]],
        [4, [597, 299, 298, 150, 149, 149, 149, 75, 75, 75, 74, 75, 74, 75, 74,
             38, 37, 38, 37, 38, 37, 37, 37, 38, 37, 37, 37, 38, 37, 37, 37]],
        [5, [139, 70, 69, 35, 35, 35, 34, 18, 17, 18, 17, 18, 17, 17, 17,
             9, 9, 9, 8, 9, 9, 9, 8, 9, 9, 9, 8, 9, 8, 9, 8,
             5, 4, 5, 4, 5, 4, 4, 4, 5, 4, 5, 4, 5, 4, 4, 4,
             5, 4, 5, 4, 5, 4, 4, 4, 5, 4, 4, 4, 5, 4, 4, 4]],
        [6, [139, 70, 69, 35, 35, 35, 34, 18, 17, 18, 17, 18, 17, 17, 17,
             9, 9, 9, 8, 9, 9, 9, 8, 9, 9, 9, 8, 9, 8, 9, 8,
             5, 4, 5, 4, 5, 4, 4, 4, 5, 4, 5, 4, 5, 4, 4, 4,
             5, 4, 5, 4, 5, 4, 4, 4, 5, 4, 4, 4, 5, 4, 4, 4,
             3, 2, 2, 2, 3, 2, 2, 2, 3, 2, 2, 2, 2, 2, 2, 2,
             3, 2, 2, 2, 3, 2, 2, 2, 3, 2, 2, 2, 2, 2, 2, 2,
             3, 2, 2, 2, 3, 2, 2, 2, 3, 2, 2, 2, 2, 2, 2, 2,
             3, 2, 2, 2, 2, 2, 2, 2, 3, 2, 2, 2, 2, 2, 2, 2]],
    ]
    uneven = [
        [2, [3264, 3264, 3261, 3255], [13044, 6528, 6516, 3264, 3264, 3261, 3255]],
        [2, [5000, 500, 1, 4499], [10000, 5500, 4500, 5000, 500, 1, 4499]],
        [3, [10, 11, 1, 9, 5, 20, 7, 6],
         [69, 31, 38, 21, 10, 25, 13, 10, 11, 1, 9, 5, 20, 7, 6]],
    ]
    return {"source": "test_constructors.c:149-254,342-436,451-470",
            "layout": "breadth-first node sizes: root, then per level left to right; "
                      "last level = diagonal leaves",
            "halves": halves, "uneven_custom": uneven}


def recompress():
    # test_hodlr_hodlr_algebra.c:1575-1669 (QR method) and 1753-1797 (large s)
    def zero_cases(ms, ss, ns):
        cases = []
        m, s, n = ms[0], ss[0], ns[0]
        cases.append({"m": m, "n": n, "s": s, "u": {"0,0": 1.0}, "v": {"0,0": 1.0},
                      "s_exp": 1, "u_exp": {"0,0": 1.0}, "v_exp": {"0,0": 1.0}})
        m, s, n = ms[1], ss[1], ns[1]
        cases.append({"m": m, "n": n, "s": s, "u": {f"{m-1},0": 1.0}, "v": {"0,0": 1.0},
                      "s_exp": 1, "u_exp": {f"{m-1},0": -1.0}, "v_exp": {"0,0": -1.0}})
        m, s, n = ms[2], ss[2], ns[2]
        cases.append({"m": m, "n": n, "s": s, "u": {"0,0": 1.0}, "v": {f"{n-1},0": 1.0},
                      "s_exp": 1, "u_exp": {"0,0": 1.0}, "v_exp": {f"{n-1},0": 1.0}})
        m, s, n = ms[3], ss[3], ns[3]
        cases.append({"m": m, "n": n, "s": s, "u": {"0,0": 1.0, "0,1": 1.0},
                      "v": {f"{n-1},0": 1.0, f"{n-1},1": 1.0},
                      "s_exp": 1, "u_exp": {"0,0": 2.0}, "v_exp": {f"{n-1},0": 1.0}})
        return cases

    def laplace_case(s):
        m = n = 11
        u = {"9,0": 2.0, "10,0": -4.0, "10,1": 1.0, "0,2": -0.5}
        v = {"0,0": -1.0, "0,1": 1.0, "1,1": -2.0, "9,2": -2.0, "10,2": 1.0}
        u_exp = {"9,0": 1.887256638321, "10,0": -5.380155478673, "0,1": 1.118033988750,
                 "9,1": -6.20633538e-17, "0,2": -3.89445209e-17,
                 "9,2": 6.62013882871e-1, "10,2": 2.32221931143e-1}
        v_exp = {"0,0": -0.943628319160, "1,0": 0.331006941436,
                 "9,1": 0.894427191000, "10,1": -0.447213595500,
                 "0,2": -0.331006941436, "1,2": -0.943628319160}
        return {"m": m, "n": n, "s": s, "u": u, "v": v, "s_exp": 3,
                "u_exp": u_exp, "v_exp": v_exp}

    return {
        "source": "test_hodlr_hodlr_algebra.c:1575-1669, 1753-1797",
        "format": "sparse entries 'row,col': value; unlisted entries are 0; thr 1e-8",
        "qr": zero_cases([10, 10, 9, 6], [2, 3, 4, 2], [10, 9, 10, 6]) + [laplace_case(3)],
        "large_s": zero_cases([10, 10, 9, 6], [10, 9, 9, 12], [10, 9, 10, 6])
                   + [laplace_case(13)],
    }


def compress_block():
    # test_constructors.c:725-822: block of a full matrix, expected factors
    return {
        "source": "test_constructors.c:725-822 (thr 1e-8)",
        "cases": [
            {"name": "L_bl", "full": "laplacian", "m_full": 10, "r0": 5, "c0": 0,
             "m": 5, "n": 5, "s": 1, "u_exp": {"0,0": 1.0}, "v_exp": {"4,0": -1.0}},
            {"name": "L_tr", "full": "laplacian", "m_full": 11, "r0": 0, "c0": 5,
             "m": 5, "n": 6, "s": 1, "u_exp": {"4,0": -1.0}, "v_exp": {"0,0": 1.0}},
            {"name": "L0.5S_bl", "full": "laplacian_corners", "m_full": 11, "r0": 5,
             "c0": 0, "m": 6, "n": 5, "s": 2,
             "u_exp": {"0,0": 1.0, "5,1": -0.5}, "v_exp": {"4,0": -1.0, "0,1": -1.0}},
            {"name": "L0.5S_tr", "full": "laplacian_corners", "m_full": 10, "r0": 0,
             "c0": 5, "m": 5, "n": 5, "s": 2,
             "u_exp": {"4,0": 1.0, "0,1": 0.5}, "v_exp": {"0,0": -1.0, "4,1": 1.0}},
            {"name": "I", "full": "identity", "m_full": 18, "r0": 5, "c0": 0, "m": 5,
             "n": 4, "s": 1, "u_exp": {}, "v_exp": {"0,0": 1.0}},
            {"name": "0", "full": "zeros", "m_full": 7, "r0": 0, "c0": 0, "m": 7, "n": 7,
             "s": 1, "u_exp": {}, "v_exp": {"0,0": 1.0}},
            {"name": "L_skew", "full": "laplacian", "m_full": 13, "r0": 9, "c0": 0,
             "m": 4, "n": 9, "s": 1, "u_exp": {"0,0": 1.0}, "v_exp": {"8,0": -1.0}},
        ],
    }


def products():
    # test_vector_algebra.c:54-154 and test_hodlr_dense_algebra.c:163-276:
    # m=21, thr 0.1, heights 1..3; vector of 10s -> sparse exact answers
    return {
        "source": "test_vector_algebra.c:54-154, test_hodlr_dense_algebra.c:163-276",
        "m": 21, "thr": 0.1, "heights": [1, 2, 3],
        "laplacian_vector_10": {
            "L": {"0": 10.0, "20": 10.0},
            "L0.5S": {"0": 15.0, "20": 15.0},
            "L0.5A": {"0": 10.0, "20": 15.0},
        },
        "laplacian_dense_10_rows": {"L": {"0": 10.0, "20": 10.0},
                                    "L0.5S": {"0": 15.0, "20": 15.0},
                                    "L0.5A": {"0": 10.0, "20": 15.0},
                                    "L0.5A_transpose": {"0": 15.0, "20": 10.0}},
        "offdiag_pair": [
            # test_vector_algebra.c:239-332: parent with m, n, s=1 blocks
            {"len": 10, "m": 4, "n": 4, "offset": 2, "u_tr": 5.0, "v_tr": 1.0,
             "u_bl": -5.0, "v_bl": 1.0, "x": 0.5, "out0": -10.0,
             "expected": "out0 everywhere; 0 on [offset, offset+m); -20 on [offset+m, offset+m+n)"},
        ],
    }


def workspace_sizes():
    # test_hodlr_dense_algebra.c:51-113
    return {"source": "test_hodlr_dense_algebra.c:51-113",
            "uniform": [[h, h, 2048, h * 2048] for h in range(1, 5)],
            "note": "height h, all ranks h, matrix_a 2048 -> h*2048; one rank raised "
                    "to 4+h at index 2h of the level-order rank list -> (4+h)*42"}


if __name__ == "__main__":
    for name, fn in (("block_sizes", block_sizes), ("recompress", recompress),
                     ("compress_block", compress_block), ("products", products),
                     ("workspace_sizes", workspace_sizes)):
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(fn(), f, indent=1)
    print("written")
