"""
Writes the known-answer fixtures used by the parity tests.

The numbers are the literal expected values and inputs held by the reference's
own unit tests (MeteoSwiss fast-barnes-py, BSD-3-Clause); each case names the
test file and lines it was taken from.  Nothing here imports or runs the
reference: the values are data.  Run from this directory:

    python make_known_answers.py
"""
import json

CASES = [
    # tests/AccumulationTest.py test_accumulate_array_1_fold :34-59
    dict(src="tests/AccumulationTest.py:40-45", fn="plain", size=32, rect_len=9, num_iter=1, alpha=0.0,
         impulses=[[10, 1.0]], expect=[[0, 6, 0.0], [6, 15, 1.0], [15, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:48-52", fn="plain", size=32, rect_len=9, num_iter=1, alpha=0.0,
         impulses=[[2, 1.0]], expect=[[0, 7, 1.0], [7, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:55-59", fn="plain", size=32, rect_len=9, num_iter=1, alpha=0.0,
         impulses=[[30, 1.0]], expect=[[0, 26, 0.0], [26, 32, 1.0]]),
    # test_accumulate_array_2_fold :61-86
    dict(src="tests/AccumulationTest.py:67-72", fn="plain", size=32, rect_len=9, num_iter=2, alpha=0.0,
         impulses=[[16, 1.0]],
         expect=[[0, 8, 0.0], [8, 25, [1, 2, 3, 4, 5, 6, 7, 8, 9, 8, 7, 6, 5, 4, 3, 2, 1]], [25, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:75-79", fn="plain", size=32, rect_len=9, num_iter=2, alpha=0.0,
         impulses=[[2, 1.0]], expect=[[0, 11, [5, 6, 7, 7, 7, 6, 5, 4, 3, 2, 1]], [11, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:82-86", fn="plain", size=32, rect_len=9, num_iter=2, alpha=0.0,
         impulses=[[30, 1.0]], expect=[[0, 22, 0.0], [22, 32, [1, 2, 3, 4, 5, 6, 6, 6, 6, 5]]]),
    # test_accumulate_tail_array_1_fold :88-113
    dict(src="tests/AccumulationTest.py:94-99", fn="tail", size=32, rect_len=7, num_iter=1, alpha=0.25,
         impulses=[[10, 1.0]], expect=[[0, 6, 0.0], [6, 15, [0.25, 1, 1, 1, 1, 1, 1, 1, 0.25]], [15, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:102-106", fn="tail", size=32, rect_len=7, num_iter=1, alpha=0.25,
         impulses=[[2, 1.0]], expect=[[0, 7, [1, 1, 1, 1, 1, 1, 0.25]], [7, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:109-113", fn="tail", size=32, rect_len=7, num_iter=1, alpha=0.25,
         impulses=[[30, 1.0]], expect=[[0, 26, 0.0], [26, 32, [0.25, 1, 1, 1, 1, 1]]]),
    # test_accumulate_tail_array_2_fold :115-140
    dict(src="tests/AccumulationTest.py:121-126", fn="tail", size=32, rect_len=7, num_iter=2, alpha=0.5,
         impulses=[[16, 1.0]],
         expect=[[0, 8, 0.0], [8, 25, [0.25, 1, 2, 3, 4, 5, 6, 7, 7.5, 7, 6, 5, 4, 3, 2, 1, 0.25]], [25, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:129-133", fn="tail", size=32, rect_len=7, num_iter=2, alpha=0.5,
         impulses=[[2, 1.0]], expect=[[0, 11, [4.5, 5.5, 6.25, 6.5, 6, 5, 4, 3, 2, 1, 0.25]], [11, 32, 0.0]]),
    dict(src="tests/AccumulationTest.py:136-140", fn="tail", size=32, rect_len=7, num_iter=2, alpha=0.5,
         impulses=[[30, 1.0]], expect=[[0, 22, 0.0], [22, 32, [0.25, 1, 2, 3, 4, 5, 5.5, 5.5, 5.25, 4.5]]]),
]

# test_accumulate_array_versions :142-165 -- equivalence pairs on one input
VERSIONS = dict(src="tests/AccumulationTest.py:150-165", size=32,
                impulses=[[3, 1.0], [9, 2.5], [21, -1.25], [30, 1.0]], num_iter=3,
                pairs=[dict(plain_rect_len=9, tail_rect_len=9, alpha=0.0),
                       dict(plain_rect_len=9, tail_rect_len=7, alpha=1.0)])

# tests/GaussianApproximationTest.py:35-37 and :69-73 -- tolerance tables
GAUSS = dict(src="tests/GaussianApproximationTest.py:35-100", sigmas=[0.6, 0.8, 1.0, 1.5, 2.0], resol=512.0,
             width_sigmas=3.5,
             plain=[[3, 0.02395], [4, 0.01405], [5, 0.01232], [6, 0.01004]],
             opt=[[3, 0.02395], [4, 0.01405], [5, 0.012325], [6, 0.010045]])

# tests/BasicTest.py:42-101 -- inputs of the smoke/range tests
BASIC = dict(
    src="tests/BasicTest.py:44-99",
    d1=dict(step=1.0 / 64, size=129,
            pts=[0.15, 0.2, 0.33, 0.35, 0.46, 0.59, 0.61, 0.66, 0.83, 0.98, 1.21, 1.29, 1.4, 1.57, 1.6, 1.79],
            val=[1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 1.0],
            sigmas=[0.5, 0.2, 0.05], num_iters=[1, 2, 3, 4, 5, 6, 8, 10], eps=1e-15),
    d2=dict(step=1.0 / 64, size=[129, 129],
            data=[[0.15, 0.07, 1.0], [0.22, 0.11, 0.0], [0.49, 0.21, 1.0], [0.63, 0.09, 0.0], [0.94, 0.20, 1.0],
                  [1.02, 0.18, 1.0], [1.11, 0.04, 0.0], [1.38, 0.18, 0.0], [1.74, 0.24, 1.0], [1.87, 0.21, 1.0],
                  [0.10, 0.47, 0.0], [0.31, 0.51, 0.0], [0.40, 0.31, 1.0], [0.70, 0.49, 0.0], [0.91, 0.50, 1.0],
                  [1.16, 0.38, 1.0], [1.19, 0.34, 0.0], [1.35, 0.48, 0.0], [1.69, 0.46, 1.0], [1.80, 0.36, 0.0],
                  [0.04, 0.67, 1.0], [0.37, 0.81, 1.0], [0.59, 0.71, 1.0], [0.88, 0.59, 0.0], [0.93, 0.81, 1.0],
                  [1.01, 0.88, 1.0], [1.26, 0.84, 0.0], [1.44, 0.78, 0.0], [1.52, 0.86, 1.0], [1.91, 0.66, 0.0],
                  [0.14, 0.99, 0.0], [0.17, 1.05, 1.0], [0.38, 1.21, 1.0], [0.43, 0.96, 0.0], [0.83, 1.11, 1.0],
                  [1.00, 1.00, 1.0], [1.13, 0.98, 0.0], [1.24, 1.23, 1.0], [1.48, 1.16, 1.0], [1.60, 1.08, 0.0],
                  [0.34, 1.28, 1.0], [0.37, 1.47, 1.0], [0.68, 1.37, 0.0], [0.93, 1.52, 0.0], [0.97, 1.39, 1.0],
                  [1.03, 1.50, 1.0], [1.44, 1.42, 0.0], [1.54, 1.51, 1.0], [1.78, 1.49, 0.0], [1.89, 1.60, 1.0],
                  [0.17, 1.79, 1.0], [0.31, 1.95, 0.0], [0.55, 1.63, 0.0], [0.85, 1.78, 1.0], [0.92, 1.65, 1.0],
                  [1.07, 1.90, 0.0], [1.30, 1.82, 1.0], [1.46, 1.71, 1.0], [1.73, 1.69, 0.0], [1.91, 1.91, 1.0]],
            sigmas=[0.5, 0.2, 0.05], num_iters=[1, 2, 3, 4, 5, 6, 8, 10], eps=5e-13),
    d3=dict(step=1.0 / 64, size=[129, 129, 129], num_values=307, sigmas=[0.5, 0.2, 0.05],
            num_iters=[1, 3, 4, 6, 10], eps=1e-12),
    kernel_check=dict(src="tests/BasicTest.py:105-148", sigma=0.5, step=1.0 / 32, num_iters=[3, 4, 6],
                      pts=[0.2, 0.4, 0.7], val=[18.5, 17.0, 19.25]),
)

if __name__ == '__main__':
    with open('accumulation_known_answers.json', 'w') as f:
        json.dump(dict(license="known-answer data from MeteoSwiss fast-barnes-py tests (BSD-3-Clause)",
                       accumulation=CASES, accumulation_versions=VERSIONS, gaussian=GAUSS, basic=BASIC),
                  f, indent=1)
    print('wrote accumulation_known_answers.json')
