// voce_serve.cpp -- request server (placeholder translation unit; filled in below)
