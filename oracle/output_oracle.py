"""CPU oracle for the CSV output format (TEST INFRASTRUCTURE ONLY).

Restates save_matrix_csv (tsw_rtm/visualizer.py:19-39): a header of
"row_name/col_name" and the column axis formatted '.6g', then one line per
row with the row-axis value ('.6g') and the values ('.6e'), comma separated.
"""


def save_matrix_csv(path, matrix, row_axis, col_axis, row_name="row", col_name="col"):
    with open(path, "w") as fh:
        fh.write(f"{row_name}/{col_name}," + ",".join(format(v, ".6g") for v in col_axis) + "\n")
        for i in range(len(row_axis)):
            fh.write(format(row_axis[i], ".6g") + "," +
                     ",".join(format(v, ".6e") for v in matrix[i]) + "\n")
