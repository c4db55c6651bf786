"""Checkpoint / resume, float16 result quantization and the caustica-result/1 file layout."""
