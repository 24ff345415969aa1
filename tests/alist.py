"""MacKay alist reader (the format of crates/com-fec/src/ldpc.rs:300-362): rows of H in file order."""


def read_alist(path):
    nums = []
    for line in open(path):
        nums.extend(int(t) for t in line.split("#")[0].split())
    n, m, _, _ = nums[:4]
    pos = 4
    col_deg = nums[pos:pos + n]
    pos += n
    row_deg = nums[pos:pos + m]
    pos += m
    cols = []
    for d in col_deg:
        # lists may be padded with zeros up to the largest degree
        lst = []
        while len(lst) < d:
            v = nums[pos]
            pos += 1
            if v:
                lst.append(v - 1)
        while pos < len(nums) and nums[pos] == 0:
            pos += 1
        cols.append(lst)
    rows = []
    for d in row_deg:
        lst = []
        while len(lst) < d:
            v = nums[pos]
            pos += 1
            if v:
                lst.append(v - 1)
        while pos < len(nums) and nums[pos] == 0 and len(rows) < m - 1:
            pos += 1
        rows.append(lst)
    # the two descriptions must agree
    from_cols = sorted((r, c) for c, lst in enumerate(cols) for r in lst)
    from_rows = sorted((r, c) for r, lst in enumerate(rows) for c in lst)
    assert from_cols == from_rows, "alist column and row lists disagree"
    return n, rows
