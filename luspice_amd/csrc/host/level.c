/* level.c -- column dependency levels with the reference's C API
 * (include/level.h; behaviour of src/level.c:20-48 of the reference).
 *
 * level[k] = max over the strictly-upper entries j of U(:,k) (positions,
 * diagonal stored last) of level[j] + 1, or 0; nlevels = max level + 1.
 * Columns of one level have no dependency on each other.
 */
#include <string.h>

#include "level.h"

int dsh_level_compute(int n, const int *Up, const int *Ui, int *level_out, int *nlevels_out)
{
    int deepest = 0;
    for (int k = 0; k < n; k++) {
        int lv = 0;
        for (int p = Up[k]; p < Up[k + 1] - 1; p++) {
            const int j = Ui[p];
            if (j >= 0 && j < k && level_out[j] >= lv) lv = level_out[j] + 1;
        }
        level_out[k] = lv;
        if (lv > deepest) deepest = lv;
    }
    *nlevels_out = deepest + 1;
    return 0;
}

int dsh_level_hist(int n, const int *level, int nlevels, int *hist_out)
{
    memset(hist_out, 0, (size_t)nlevels * sizeof(int));
    for (int k = 0; k < n; k++)
        if (level[k] >= 0 && level[k] < nlevels) hist_out[level[k]]++;
    return 0;
}
