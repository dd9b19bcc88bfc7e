/*
 * o_parse.c -- oracle restatement of the reference's parameter parsing.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   ECM reader ............ src/ECM.ml:18-72
 *   Newick lexer/parser ... lib/CamlPaml/NewickLexer.mll:5-14, NewickParser.mly:7-24
 *   Newick.subtree ........ lib/CamlPaml/Newick.ml:34-41
 *   Newick.total_length ... lib/CamlPaml/Newick.ml:52-61
 *   T.of_newick ........... lib/CamlPaml/T.ml:57-112
 */
#include "oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(char *err, int errlen, int code, const char *fmt, ...) {
  if (err && errlen > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
  }
  return code;
}

/* OCaml float_of_string accepts what strtod accepts, but the whole token
 * must be consumed. */
static int parse_float_tok(const char *s, size_t len, double *out) {
  char buf[128];
  if (len == 0 || len >= sizeof buf) return -1;
  memcpy(buf, s, len);
  buf[len] = 0;
  char *end = NULL;
  errno = 0;
  double v = strtod(buf, &end);
  if (end != buf + len) return -1;
  *out = v;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* ECM.import_parameters (src/ECM.ml:18-72)                                 */

static const char NUC[4] = {'A', 'C', 'G', 'T'};

static int dna_index(char c) { /* Code.DNA.index (lib/CamlPaml/Code.ml:25-30) */
  switch (c) {
  case 'A': case 'a': return 0;
  case 'C': case 'c': return 1;
  case 'G': case 'g': return 2;
  case 'T': case 't': return 3;
  default: return -1;
  }
}

/* split a line on single spaces (Str.split re_sp), trim each piece and keep
 * the non-empty ones; returns the count (max cap) */
static int split_floats(const char *line, size_t len, double *vals, int cap, int *bad) {
  int n = 0;
  size_t i = 0;
  *bad = 0;
  while (i <= len) {
    size_t j = i;
    while (j < len && line[j] != ' ') j++;
    size_t a = i, b = j;
    while (a < b && isspace((unsigned char)line[a])) a++;
    while (b > a && isspace((unsigned char)line[b - 1])) b--;
    if (b > a) {
      double v;
      if (parse_float_tok(line + a, b - a, &v) != 0) {
        *bad = 1;
        return n;
      }
      if (n < cap) vals[n] = v;
      n++;
    }
    i = j + 1;
  }
  return n;
}

int o_ecm_read(const char *path, double *S, double *pi, char *err, int errlen) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail(err, errlen, O_EFAIL, "cannot open %s", path);
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *txt = (char *)malloc((size_t)sz + 1);
  if (fread(txt, 1, (size_t)sz, f) != (size_t)sz) {
    fclose(f);
    free(txt);
    return fail(err, errlen, O_EFAIL, "read error %s", path);
  }
  fclose(f);
  txt[sz] = 0;
  /* lines (File.lines_of drops the '\n') */
  int cap = 256, nlines = 0;
  const char **ls = (const char **)malloc(sizeof(char *) * (size_t)cap);
  size_t *ll = (size_t *)malloc(sizeof(size_t) * (size_t)cap);
  char *p = txt;
  while (*p) {
    char *e = strchr(p, '\n');
    size_t len = e ? (size_t)(e - p) : strlen(p);
    if (nlines == cap) {
      cap *= 2;
      ls = (const char **)realloc(ls, sizeof(char *) * (size_t)cap);
      ll = (size_t *)realloc(ll, sizeof(size_t) * (size_t)cap);
    }
    ls[nlines] = p;
    ll[nlines] = len;
    nlines++;
    if (!e) break;
    p = e + 1;
  }
  int rc = 0;
  double row[80];
  double raw[63][64];
  if (nlines < 71) {
    rc = fail(err, errlen, O_EINVAL, "ECM.import_parameters: too few lines (%d)", nlines);
    goto out;
  }
  for (int r = 0; r < 63; r++) {
    int bad;
    int cnt = split_floats(ls[r], ll[r], row, 80, &bad);
    if (bad) {
      rc = fail(err, errlen, O_EFAIL, "float_of_string on ECM line %d", r + 1);
      goto out;
    }
    if (cnt < r + 1) {
      rc = fail(err, errlen, O_EINVAL, "index out of bounds (ECM line %d)", r + 1);
      goto out;
    }
    for (int j = 0; j <= r; j++) raw[r][j] = row[j];
  }
  for (int i = 0; i < 64; i++)
    for (int j = 0; j < 64; j++) {
      double v;
      if (i == 0 && j == 0) v = 0.;
      else if (i > j) v = raw[i - 1][j];
      else if (i < j) v = raw[j - 1][i];
      else v = 0.;
      S[i * 64 + j] = v;
    }
  {
    size_t a = 0, b = ll[63];
    while (a < b && isspace((unsigned char)ls[63][a])) a++;
    if (a != b) {
      rc = fail(err, errlen, O_EFAIL, "ECM.import_parameters");
      goto out;
    }
  }
  {
    int bad;
    int cnt = split_floats(ls[64], ll[64], row, 80, &bad);
    if (bad) {
      rc = fail(err, errlen, O_EFAIL, "float_of_string on ECM frequency line");
      goto out;
    }
    if (cnt != 64) {
      rc = fail(err, errlen, O_EINVAL, "ECM frequency line has %d entries", cnt);
      goto out;
    }
    for (int j = 0; j < 64; j++) pi[j] = row[j];
  }
  {
    int ncod = 0;
    for (int li = 67; li <= 70; li++) {
      const char *s = ls[li];
      size_t len = ll[li], i = 0;
      while (i <= len) {
        size_t j = i;
        while (j < len && s[j] != ' ') j++;
        size_t a = i, b = j;
        while (a < b && isspace((unsigned char)s[a])) a++;
        while (b > a && isspace((unsigned char)s[b - 1])) b--;
        if (b > a) {
          if (b - a != 3) {
            rc = fail(err, errlen, O_EFAIL, "Assert_failure: codon length");
            goto out;
          }
          int i0 = dna_index(s[a]), i1 = dna_index(s[a + 1]), i2 = dna_index(s[a + 2]);
          if (i0 < 0 || i1 < 0 || i2 < 0) {
            rc = fail(err, errlen, O_EINVAL, "unrecognized nucleotide");
            goto out;
          }
          if (16 * i0 + 4 * i1 + i2 != ncod) {
            rc = fail(err, errlen, O_EFAIL, "ECM.import_parameters: incorrect codon order");
            goto out;
          }
          ncod++;
        }
        i = j + 1;
      }
    }
    if (ncod != 64) {
      rc = fail(err, errlen, O_EFAIL, "ECM.import_parameters: incorrect codon order");
      goto out;
    }
  }
  (void)NUC;
out:
  free(ls);
  free(ll);
  free(txt);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Newick (NewickLexer.mll, NewickParser.mly, Newick.ml)                    */

typedef struct nw {
  struct nw **ch;
  int nch;
  char *lbl;
  int has_bl;
  double bl;
} nw;

static void nw_free(nw *x) {
  if (!x) return;
  for (int i = 0; i < x->nch; i++) nw_free(x->ch[i]);
  free(x->ch);
  free(x->lbl);
  free(x);
}

enum { T_LP, T_RP, T_COMMA, T_COLON, T_BL, T_LABEL, T_EOF, T_ERR };

typedef struct {
  const char *s;
  size_t pos, len;
  int tok;
  const char *tstart;
  size_t tlen;
  double bl;
} lexer;

static int is_lblch(char c) { return isalnum((unsigned char)c) || c == '_' || c == '.'; }

/* token rule (NewickLexer.mll:5-14): whitespace and ';' are skipped; the
 * longest match wins and BRANCHLEN wins ties with LABEL. */
static void lex_next(lexer *L) {
  while (L->pos < L->len) {
    char c = L->s[L->pos];
    if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == ';') L->pos++;
    else break;
  }
  if (L->pos >= L->len) {
    L->tok = T_EOF;
    return;
  }
  char c = L->s[L->pos];
  if (c == '(') { L->tok = T_LP; L->pos++; return; }
  if (c == ')') { L->tok = T_RP; L->pos++; return; }
  if (c == ',') { L->tok = T_COMMA; L->pos++; return; }
  if (c == ':') { L->tok = T_COLON; L->pos++; return; }
  if (is_lblch(c)) {
    size_t a = L->pos, b = a;
    int allnum = 1;
    while (b < L->len && is_lblch(L->s[b])) {
      if (!(isdigit((unsigned char)L->s[b]) || L->s[b] == '.')) allnum = 0;
      b++;
    }
    /* BRANCHLEN matches the longest [0-9.]+ prefix; LABEL the whole run.
     * LABEL is chosen only when strictly longer. */
    size_t nb = a;
    while (nb < L->len && (isdigit((unsigned char)L->s[nb]) || L->s[nb] == '.')) nb++;
    (void)allnum;
    if (nb == b) {
      L->tstart = L->s + a;
      L->tlen = nb - a;
      if (parse_float_tok(L->tstart, L->tlen, &L->bl) != 0) {
        L->tok = T_ERR;
        return;
      }
      L->tok = T_BL;
      L->pos = nb;
      return;
    }
    L->tstart = L->s + a;
    L->tlen = b - a;
    L->tok = T_LABEL;
    L->pos = b;
    return;
  }
  L->tok = T_ERR;
}

static nw *nw_new(void) { return (nw *)calloc(1, sizeof(nw)); }

/* label: LABEL | LABEL ':' BRANCHLEN | ':' BRANCHLEN  (NewickParser.mly:17-20) */
static int parse_label(lexer *L, nw *x) {
  if (L->tok == T_LABEL) {
    x->lbl = strndup(L->tstart, L->tlen);
    lex_next(L);
    if (L->tok == T_COLON) {
      lex_next(L);
      if (L->tok != T_BL) return -1;
      x->has_bl = 1;
      x->bl = L->bl;
      lex_next(L);
    }
    return 0;
  }
  if (L->tok == T_COLON) {
    x->lbl = strdup("");
    lex_next(L);
    if (L->tok != T_BL) return -1;
    x->has_bl = 1;
    x->bl = L->bl;
    lex_next(L);
    return 0;
  }
  return -1;
}

/* node: label | '(' subtrees ')' label | '(' subtrees ')'  (NewickParser.mly:12-16) */
static nw *parse_node(lexer *L, int depth) {
  if (depth > 100000) return NULL;
  nw *x = nw_new();
  if (L->tok == T_LP) {
    lex_next(L);
    int cap = 2;
    x->ch = (nw **)malloc(sizeof(nw *) * (size_t)cap);
    for (;;) {
      nw *c = parse_node(L, depth + 1);
      if (!c) {
        nw_free(x);
        return NULL;
      }
      if (x->nch == cap) {
        cap *= 2;
        x->ch = (nw **)realloc(x->ch, sizeof(nw *) * (size_t)cap);
      }
      x->ch[x->nch++] = c;
      if (L->tok == T_COMMA) {
        lex_next(L);
        continue;
      }
      break;
    }
    if (L->tok != T_RP) {
      nw_free(x);
      return NULL;
    }
    lex_next(L);
    if (L->tok == T_LABEL || L->tok == T_COLON) {
      if (parse_label(L, x) != 0) {
        nw_free(x);
        return NULL;
      }
    } else {
      x->lbl = strdup("");
    }
    return x;
  }
  if (parse_label(L, x) != 0) {
    nw_free(x);
    return NULL;
  }
  return x;
}

static nw *nw_parse(const char *text, char *err, int errlen) {
  lexer L = {text, 0, strlen(text), 0, NULL, 0, 0.};
  lex_next(&L);
  nw *root = parse_node(&L, 0);
  if (!root || L.tok != T_EOF) {
    nw_free(root);
    fail(err, errlen, O_EFAIL, "Parsing error (Newick)");
    return NULL;
  }
  return root;
}

static int keep_has(const char *keep_csv, const char *lbl) {
  size_t n = strlen(lbl);
  const char *p = keep_csv;
  while (*p) {
    const char *e = strchr(p, ',');
    size_t len = e ? (size_t)(e - p) : strlen(p);
    if (len == n && strncmp(p, lbl, n) == 0) return 1;
    if (!e) break;
    p = e + 1;
  }
  return 0;
}

/* Newick.subtree (lib/CamlPaml/Newick.ml:34-41); consumes x. */
static nw *nw_subtree(nw *x, const char *keep) {
  int kept_lbl = (x->lbl[0] == 0) || keep_has(keep, x->lbl);
  if (x->nch == 0) {
    if (kept_lbl) return x;
    nw_free(x);
    return NULL;
  }
  if (!kept_lbl) {
    nw_free(x);
    return NULL;
  }
  int m = 0;
  for (int i = 0; i < x->nch; i++) {
    nw *c = nw_subtree(x->ch[i], keep);
    if (c) x->ch[m++] = c;
  }
  x->nch = m;
  if (m == 0) {
    nw_free(x);
    return NULL;
  }
  if (m == 1) {
    nw *c = x->ch[0];
    /* Node (sst, snm, maybe_add bl sbl) */
    if (x->has_bl && c->has_bl) {
      c->bl = x->bl + c->bl;
    } else {
      c->has_bl = 0;
    }
    x->nch = 0;
    nw_free(x);
    return c;
  }
  return x;
}

static int nw_size(const nw *x) {
  int s = 1;
  for (int i = 0; i < x->nch; i++) s += nw_size(x->ch[i]);
  return s;
}

static int nw_leaves(const nw *x) {
  if (x->nch == 0) return 1;
  int s = 0;
  for (int i = 0; i < x->nch; i++) s += nw_leaves(x->ch[i]);
  return s;
}

typedef struct {
  int n, nl, next_leaf, next_int, bad;
  int *parent, *lc, *rc;
  double *branch;
  const nw **nodes;
} tfill;

/* T.of_newick's find_leaves + fill (lib/CamlPaml/T.ml:68-110): leaves are
 * numbered in left-to-right order, internal nodes as they complete. */
static int t_fill(tfill *F, const nw *x) {
  if (x->nch == 0) {
    int i = F->next_leaf++;
    F->nodes[i] = x;
    F->branch[i] = x->has_bl ? x->bl : NAN;
    return i;
  }
  if (x->nch != 2) {
    F->bad = 1;
    return -1;
  }
  int l = t_fill(F, x->ch[0]);
  if (F->bad) return -1;
  int r = t_fill(F, x->ch[1]);
  if (F->bad) return -1;
  int i = F->next_int++;
  F->nodes[i] = x;
  F->branch[i] = x->has_bl ? x->bl : NAN;
  F->parent[l] = i;
  F->parent[r] = i;
  F->lc[i] = l;
  F->rc[i] = r;
  return i;
}

int o_newick_tree(const char *text, const char *keep_csv, int max_n, int *n_out,
                  int *parent, int *lc, int *rc, double *branch, char *labels_buf,
                  int labels_len, char *err, int errlen) {
  nw *t = nw_parse(text, err, errlen);
  if (!t) return O_EFAIL;
  if (keep_csv) {
    t = nw_subtree(t, keep_csv);
    /* PhyloCSF.ml:431-433 */
    if (!t || nw_leaves(t) <= 1) {
      nw_free(t);
      return fail(err, errlen, O_EFAIL, "specify at least two available --species");
    }
  }
  int n = nw_size(t);
  if (n < 3 || n % 2 == 0 || n > max_n) {
    nw_free(t);
    return fail(err, errlen, O_EINVAL,
                "CamlPaml.T.of_newick: input is not a rooted, bifurcating tree");
  }
  tfill F;
  memset(&F, 0, sizeof F);
  F.n = n;
  F.nl = (n + 1) / 2;
  F.next_leaf = 0;
  F.next_int = F.nl;
  F.parent = parent;
  F.lc = lc;
  F.rc = rc;
  F.branch = branch;
  F.nodes = (const nw **)calloc((size_t)n, sizeof(nw *));
  for (int i = 0; i < n; i++) {
    parent[i] = -1;
    lc[i] = rc[i] = -1;
  }
  int root = t_fill(&F, t);
  if (F.bad || root != n - 1 || F.next_leaf != F.nl) {
    free(F.nodes);
    nw_free(t);
    return fail(err, errlen, O_EINVAL,
                "CamlPaml.T.of_newick: input is not a rooted, bifurcating tree");
  }
  if (labels_buf && labels_len > 0) {
    size_t off = 0;
    labels_buf[0] = 0;
    for (int i = 0; i < n; i++) {
      const char *l = F.nodes[i]->lbl;
      size_t ln = strlen(l);
      if (off + ln + 2 > (size_t)labels_len) break;
      memcpy(labels_buf + off, l, ln);
      off += ln;
      labels_buf[off++] = (i + 1 < n) ? '\n' : 0;
    }
    labels_buf[off < (size_t)labels_len ? off : (size_t)labels_len - 1] = 0;
  }
  free(F.nodes);
  nw_free(t);
  *n_out = n;
  return 0;
}

static int nw_total(const nw *x, int is_root, double *acc) {
  double bl;
  if (is_root) bl = 0.;
  else if (x->has_bl) bl = x->bl;
  else return -1;
  double s = 0.;
  for (int i = 0; i < x->nch; i++) {
    double c;
    if (nw_total(x->ch[i], 0, &c) != 0) return -1;
    s += c;
  }
  /* total_length' : bl +. fold_left (+.) 0. (children) */
  *acc = bl + s;
  return 0;
}

int o_newick_total_length(const char *text, const char *keep_csv, double *out, char *err,
                          int errlen) {
  nw *t = nw_parse(text, err, errlen);
  if (!t) return O_EFAIL;
  if (keep_csv) t = nw_subtree(t, keep_csv);
  if (!t) {
    *out = 0.;
    return 1; /* None: the CLI's bls maps it to 0 (PhyloCSF.ml:261) */
  }
  int rc = nw_total(t, 1, out);
  nw_free(t);
  if (rc) return fail(err, errlen, O_EINVAL, "CamlPaml.Newick.total_length: unspecified branch length");
  return 0;
}
