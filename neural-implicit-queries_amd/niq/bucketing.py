"""Power-of-two bucket sizes (src/bucketing.py:7-14, 35-36).

The reference pads every working array to these sizes to bound JIT
recompiles; this build keeps them only where they are visible in results
(returned array sizes and the ``N_evals`` count).  Compaction itself happens on
the device inside the tree and ray kernels.
"""
bucket_sizes = [2 ** s for s in range(7, 31)]


def get_next_bucket_size(s):
    for b in bucket_sizes:
        if s <= b:
            return b
    raise ValueError("max bucket size exceeded")


def fits_in_smaller_bucket(size, curr_bucket_size):
    return get_next_bucket_size(size) < curr_bucket_size
